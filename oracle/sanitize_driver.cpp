// Sanitizer driver for the CPU oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
// Built together with mdas_oracle.cpp under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_oracle_sanitized.py) and run as its
// own process: every oracle entry point the tests use, over small grids of
// every dimension and scene, so that an out-of-bounds access, a use after
// free or undefined arithmetic in the oracle fails the CPU suite.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mdas_oracle.h"

static int check(bool ok, const char* what) {
    if (!ok) std::fprintf(stderr, "sanitize_driver: %s failed\n", what);
    return ok ? 0 : 1;
}

int main() {
    int bad = 0;
    struct Case { int dim, kind, b, e, W, H, spp, cand; float alpha; };
    const Case cases[] = {{3, 1, 0, 3, 32, 24, 8, 4, 0.25f}, {4, 2, 1, 1, 24, 16, 6, 2, 1.0f},
                          {5, 3, 0, 2, 20, 20, 10, 8, 0.015625f}, {6, 4, 1, 0, 16, 12, 12, 4, 0.25f},
                          {3, 0, 0, 0, 8, 8, 4, 1, 0.0f}, {4, 5, 0, 1, 12, 10, 9, 3, 0.5f}};
    for (const Case& c : cases) {
        std::vector<unsigned char> params((size_t)orc_make_scene(c.kind, 7, c.dim, (float)c.W, (float)c.H, nullptr, 0));
        orc_make_scene(c.kind, 7, c.dim, (float)c.W, (float)c.H, params.data(), (long long)params.size());
        const long long target = (long long)c.spp * c.W * c.H;
        orc_grid* g = orc_create(c.dim, target * 2 + 1024, c.cand, c.b, c.e, c.alpha, (float)c.W, (float)c.H);
        bad += check(g != nullptr, "create");
        if (!g) continue;
        int passes = 0;
        bad += check(orc_run(g, params.data(), target, &passes) == 0, "run");
        bad += check(orc_validate(g) == 1, "validate");
        std::vector<float> px((size_t)4 * c.W * c.H), den((size_t)4 * c.W * c.H);
        bad += check(orc_integrate(g, px.data(), c.W, c.H) == 0, "integrate");
        std::vector<unsigned char> bytes((size_t)4 * c.W * c.H);
        orc_convert_to_bytes(px.data(), bytes.data(), (long long)c.W * c.H);
        orc_sampling_density(g, den.data(), c.W, c.H);
        const long long nk = orc_recon_keys(g, c.W, c.H, 8, 8, nullptr, 0);
        std::vector<uint64_t> keys((size_t)nk);
        orc_recon_keys(g, c.W, c.H, 8, 8, keys.data(), nk);
        const long long N = orc_num_samples(g), L = orc_num_leaves(g);
        std::vector<float> coords((size_t)N * c.dim), values((size_t)N * 3), box((size_t)L * 2 * c.dim),
            err((size_t)L);
        std::vector<int> idx((size_t)L * 4);
        std::vector<uint32_t> seed((size_t)L);
        orc_get_state(g, coords.data(), values.data(), idx.data(), box.data(), err.data(), seed.data());
        // a standalone pass on the state set back, then the same run again
        bad += check(orc_set_state(g, N, L, coords.data(), values.data(), idx.data(), box.data(), err.data(),
                                   seed.data(), orc_max_error(g)) == 0, "set_state");
        orc_compute_errors(g);
        const int rc = orc_adaptive_sampling(g);
        bad += check(rc == 0 || rc == -2, "adaptive_sampling");
        if (rc == 0) orc_eval_range(g, params.data(), orc_num_samples(g) - orc_new_samples(g), orc_new_samples(g));
        std::vector<double> c64((size_t)64 * c.dim), v64((size_t)64 * 3);
        for (size_t k = 0; k < c64.size(); ++k) c64[k] = coords[k % coords.size()];
        orc_eval_scene_f64(params.data(), c64.data(), c.dim, 64, v64.data());
        orc_eval_scene_grouped(params.data(), coords.data(), c.dim, N < 256 ? N : 256, 0, values.data());
        std::vector<float> q((size_t)4 * c.W * c.H);
        orc_qmc_image(params.data(), c.W, c.H, 2, 0, 0, 0, q.data());
        orc_destroy(g);
    }
    std::printf("sanitize_driver: %s\n", bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}
