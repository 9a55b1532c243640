/* mlp.h - thermal-synchrotron surrogate: shared definitions (device and host).

   Network 3 -> 64 -> 64 -> 64 -> 2 with SiLU on the hidden layers. Parameters are one FP32 vector
   (8706 floats): W1[3][64], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64][2], b4[2], with
   W[k][j] multiplying input k into output j.

   Inputs (the training box of the surrogate):
     u0 = log10(nu / nu_c) in [-4, 6],  Theta_e in [0.5, 100] (log-uniform),  theta_B in (0, pi).
   The network sees them through fixed coordinates in which the coefficients are close to
   separable, s = (ln X, ln Theta_e, ln sin theta_B) with X = nu / nu_s = 10^u0 / ((2/9) Theta^2
   sin theta_B), each mapped linearly onto [-1, 1] over its range:
     ln X in [LX_LO, LX_HI] (clamped: above LX_HI both targets sit at their floors),
     ln Theta in [ln 0.5, ln 100],  ln sin theta_B in [ln 1e-3, 0] (sin theta_B >= 1e-3 as in the
     tracer's flow model).
   Outputs (normalised by per-output mean and scale):
     y0 ~ ln f_j,  f_j = j / (n_e e^2 nu_c / c)
     y1 ~ ln f_a,  f_a = alpha / (n_e e^2 / (m_e c nu_c)) = f_j / (2 Theta (nu/nu_c)^2)
   (the Rayleigh-Jeans form of Kirchhoff's law; the exact B_nu differs by O(h nu / k T_e), below
   1e-8 at 230 GHz for T_e > 1e9 K). Targets below the floors are clamped to them: there the
   coefficients are more than e^-40 below the peak and carry no flux. */
#pragma once

namespace bhmlp {

constexpr int NIN = 3, NH = 64, NOUT = 2;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH, OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH * NOUT,
              NPARAM = OFF_B4 + NOUT;
static_assert(NPARAM == 8706, "parameter count");

constexpr double U0_LO = -4.0, U0_HI = 6.0;
constexpr double TH_LO = 0.5, TH_HI = 100.0;
constexpr double TB_LO = 0.0, TB_HI = 3.14159265358979323846;
constexpr double FLOOR_J = -40.0, FLOOR_A = -70.0;
// ln(sqrt(2) pi / 3 * 2 / 9)
constexpr double LN_FJ_CONST = -1.1113862093150109;
// surrogate coordinates s = (ln X, ln Theta, ln sin theta_B): box
constexpr double LX_LO = -17.0, LX_HI = 14.0;
constexpr double LTH_LO = -0.6931471805599453, LTH_HI = 4.605170185988092;   // ln 0.5, ln 100
constexpr double LS_LO = -6.907755278982137, LS_HI = 0.0;                    // ln 1e-3, ln 1
constexpr double LN_2_9 = -1.5040773967762742;                              // ln(2/9)
constexpr double LN10 = 2.302585092994046;

struct Norm {
    float x_c[3], x_h[3];      // input centre and half-width per surrogate coordinate
    float y_mu[2], y_sd[2];    // output normalisation
};
constexpr int NORM_FLOATS = 10;   // x_c[3], x_h[3], y_mu[2], y_sd[2] (bh_mlp_get / bh_mlp_set order)

// Surrogate deployment: the trained network tabulated over its box in the surrogate coordinates,
// plus per-cell trust flags set by the exact routine (flagged cells fall back to the exact formula
// in the tracer).
struct SurrogateTable {
    const float* tab;           // [n2][n1][n0] pairs (ln f_j, ln f_a)
    const unsigned char* bad;   // [(n2-1)][(n1-1)][(n0-1)] cells that failed the exact-routine judge
    int n0, n1, n2;
    float lo0, lo1, lo2, inv_d0, inv_d1, inv_d2, hi0;
};

}  // namespace bhmlp
