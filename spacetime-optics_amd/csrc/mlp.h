/* mlp.h - thermal-synchrotron surrogate: shared definitions (device and host).

   Network 3 -> 64 -> 64 -> 64 -> 2 with SiLU on the hidden layers. Parameters are one FP32 vector
   (8706 floats): W1[3][64], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64][2], b4[2], with
   W[k][j] multiplying input k into output j.

   Inputs (the training box of the surrogate):
     u0 = log10(nu / nu_c) in [-4, 6],  Theta_e in [0.5, 100] (log-uniform),  theta_B in (0, pi);
   normalised to x = ((u0 - 1) / 5, (log10 Theta - c1) / h1, theta_B / (pi/2) - 1).
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

struct Norm {
    float x_c1, x_h1;          // log10 Theta centre and half-width
    float y_mu[2], y_sd[2];    // output normalisation
};

// Surrogate deployment: the trained network tabulated over its box, plus per-cell trust flags set by
// the exact routine (flagged cells fall back to the exact formula in the tracer).
struct SurrogateTable {
    const float* tab;      // interleaved pairs     // [n2][n1][n0][2] (ln f_j, ln f_a)
    const unsigned char* bad;   // [(n2-1)][(n1-1)][(n0-1)] cells that failed the exact-routine judge
    int n0, n1, n2;
    float lo0, lo1, lo2, inv_d0, inv_d1, inv_d2;
};

}  // namespace bhmlp
