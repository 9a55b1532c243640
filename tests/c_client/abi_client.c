/* A plain C99 client of the C ABI (include/bhrt.h, include/bhrt_host.h): built by the test with a C
   compiler against libbhrt.so, as a C or cgo caller would. It prints the struct sizes this
   compiler lays out (one per line, in bhh_abi_sizes' order) and checks the argument errors that
   need no device. Exit status: the number of failed checks. */
#include <stdio.h>
#include <string.h>

#include "bhrt.h"
#include "bhrt_host.h"

static int fails = 0;

static void expect(int got, int want, const char* what) {
    if (got != want) {
        printf("FAIL %s: %d, want %d\n", what, got, want);
        ++fails;
    }
}

int main(void) {
    const size_t sizes[] = {sizeof(bh_scene),      sizeof(bh_tables),        sizeof(bh_frame),
                            sizeof(bh_settings),   sizeof(bh_trace_opts),    sizeof(bh_stats),
                            sizeof(bh_ray_out),    sizeof(bh_batch_frame),   sizeof(bh_batch_cfg),
                            sizeof(bh_mlp_train_cfg), sizeof(bh_mlp_report), sizeof(bh_mlp_bake_report),
                            sizeof(bh_post_params), sizeof(bh_tone_params),  sizeof(bh_hud_glyph),
                            sizeof(bh_hud_rect),   sizeof(bh_font)};
    size_t i;
    for (i = 0; i < sizeof sizes / sizeof sizes[0]; ++i) printf("%zu\n", sizes[i]);

    bh_frame frame;
    bh_trace_opts opts;
    bh_stats stats;
    bh_scene scene;
    float px[4];
    double jy = 0.0, x = 1.0, y = 1.0, out = 0.0;
    memset(&frame, 0, sizeof frame);
    memset(&opts, 0, sizeof opts);
    memset(&scene, 0, sizeof scene);
    expect(bh_render(NULL, &frame, &opts, px, &stats), BH_ERR_ARG, "bh_render without a context");
    expect(bh_update_scene(NULL, &scene), BH_ERR_ARG, "bh_update_scene without a context");
    expect(bh_radio_flux(NULL, &frame, 16, 1.0, &jy), BH_ERR_ARG, "bh_radio_flux without a context");
    expect(bh_trace_batch(NULL, NULL, 0, &opts, NULL, px, &stats), BH_ERR_ARG, "bh_trace_batch without a context");
    expect(bh_math_eval(99, &x, &y, &out, 1), BH_ERR_ARG, "bh_math_eval with an unknown function");
    expect(bh_math_eval(0, &x, &y, &out, 0), BH_OK, "bh_math_eval of nothing");
    bh_destroy(NULL);   /* a no-op */
    return fails;
}
