/*
 * bsx_capi_check.c — drives libbsx through include/bsx.h from C (no Python):
 * particle_reconstruct's lifecycle (bsx_recon_create / insert / finish /
 * destroy) and central-section extraction (bsx_map_transform, bsx_extract),
 * each compared with the FP64 CPU oracle (oracle/liboracle64.so, the checker)
 * on the same seeded inputs, plus the argument-error paths.
 *
 * Prints one line per check and exits 0 only when every check passes.
 * Run by tests/test_gpu_capi.py on a ROCm device.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "../../include/bsx.h"
#include "../../oracle/bsx_oracle.h"

static int failures = 0;

#define CHECK(cond, ...)                                                    \
    do {                                                                    \
        if (!(cond)) {                                                      \
            printf("FAIL %s:%d: ", __FILE__, __LINE__);                     \
            printf(__VA_ARGS__);                                            \
            printf("\n");                                                   \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

#define BSX(call)                                                           \
    do {                                                                    \
        int _r = (call);                                                    \
        if (_r != BSX_OK) {                                                 \
            printf("FAIL %s -> %d: %s\n", #call, _r, bsx_last_error());     \
            return 1;                                                       \
        }                                                                   \
    } while (0)

#define HIPC(call)                                                          \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) {                                             \
            printf("FAIL %s: %s\n", #call, hipGetErrorString(_e));          \
            return 1;                                                       \
        }                                                                   \
    } while (0)

/* splitmix64: a seeded generator, so the inputs are the same on every run */
static uint64_t rng_state = 0x5eed;
static double urand(void) {
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static double max_abs(const double* a, long n) {
    double m = 0;
    for (long i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
    return m;
}

/* Particles of `box` pixels: three Gaussian spots and uniform noise, random
 * views, origins near the centre, defocus 1-3 um with astigmatism. */
static void make_particles(int box, int n, float* imgs, bsx_particle* parts) {
    for (int p = 0; p < n; ++p) {
        double cx[3], cy[3];
        for (int b = 0; b < 3; ++b) {
            cx[b] = box * (0.3 + 0.4 * urand());
            cy[b] = box * (0.3 + 0.4 * urand());
        }
        for (int y = 0; y < box; ++y)
            for (int x = 0; x < box; ++x) {
                double v = urand() - 0.5;
                for (int b = 0; b < 3; ++b)
                    v += 2.0 * exp(-((x - cx[b]) * (x - cx[b]) + (y - cy[b]) * (y - cy[b])) / 8.0);
                imgs[((long)p * box + y) * box + x] = (float)v;
            }
        bsx_particle* q = &parts[p];
        memset(q, 0, sizeof(*q));
        double vx = 2 * urand() - 1, vy = 2 * urand() - 1, vz = 2 * urand() - 1;
        double r = sqrt(vx * vx + vy * vy + vz * vz) + 1e-9;
        q->view[0] = vx / r;
        q->view[1] = vy / r;
        q->view[2] = vz / r;
        q->view[3] = (2 * urand() - 1) * M_PI;
        q->ori[0] = box / 2 + (urand() - 0.5);
        q->ori[1] = box / 2 + (urand() - 0.5);
        q->hdr_ori[0] = box / 2;
        q->hdr_ori[1] = box / 2;
        q->pixel_size = 1.5;
        q->mag = 1;
        q->volt = 300e3;
        q->cs = 2.7e7;
        q->amp_shift = asin(0.07);
        q->def_avg = 1e4 + 2e4 * urand();
        q->def_dev = 1e3 * urand();
        q->ast_angle = M_PI * (urand() - 0.5);
        q->sel = 1;
        q->half = p % 2;
    }
}

static int check_reconstruction(void) {
    const int N = 32, box = 32, n = 48;
    const int ft = bsx_part_ft_size(box, 1.0, 2);
    CHECK(ft == 64, "bsx_part_ft_size(32, 1, 2) = %d, expected 64", ft);
    const long n3 = (long)N * N * N;
    float* imgs = malloc(sizeof(float) * n * box * box);
    bsx_particle* parts = malloc(sizeof(bsx_particle) * n);
    make_particles(box, n, imgs, parts);

    bsx_recon_params rp;
    memset(&rp, 0, sizeof(rp));
    rp.size = N;
    rp.box = box;
    rp.ft_size = ft;
    rp.interp = BSX_INTERP_NEAREST;
    rp.sampling = 1.5;
    rp.ctf_action = BSX_CTF_FLIP;
    rp.nhalves = 2;

    /* GPU: create, insert (device images), finish the full map and half 1 */
    float *d_imgs, *d_map;
    HIPC(hipMalloc((void**)&d_imgs, sizeof(float) * n * box * box));
    HIPC(hipMalloc((void**)&d_map, sizeof(float) * n3));
    HIPC(hipMemcpy(d_imgs, imgs, sizeof(float) * n * box * box, hipMemcpyHostToDevice));
    bsx_recon* h = NULL;
    BSX(bsx_recon_create(&h, &rp));
    BSX(bsx_recon_insert(h, d_imgs, parts, n, NULL));
    float* gmap = malloc(sizeof(float) * n3);
    double* omap = malloc(sizeof(double) * n3);
    for (int imap = 0; imap < 2; ++imap) {
        const int nmaps = imap == 0 ? 1 : 2;     /* full map, then half 1 of 2 */
        bsx_stats gs;
        BSX(bsx_recon_finish(h, imap == 0 ? 0 : 1, nmaps, d_map, NULL, NULL, &gs, NULL));
        HIPC(hipMemcpy(gmap, d_map, sizeof(float) * n3, hipMemcpyDeviceToHost));

        /* oracle: two partial maps (the halves), the same finish */
        or_recon* o = or_recon_new(&rp, 2);
        CHECK(o != NULL, "or_recon_new");
        if (!o) return 1;
        CHECK(or_recon_insert(o, parts, imgs, n, 1) == 0, "or_recon_insert");
        bsx_stats os;
        CHECK(or_recon_finish(o, imap == 0 ? 0 : 1, nmaps, omap, NULL, NULL, &os) == 0,
              "or_recon_finish");
        or_recon_free(o);
        double err = 0;
        for (long i = 0; i < n3; ++i) err = fmax(err, fabs(gmap[i] - omap[i]));
        err /= max_abs(omap, n3);
        printf("reconstruction imap %d nmaps %d: %lld particles, coverage %lld (oracle %lld), "
               "max rel voxel error %.3g\n", imap == 0 ? 0 : 1, nmaps, (long long)gs.particles,
               (long long)gs.coverage, (long long)os.coverage, err);
        CHECK(err <= 1e-4, "map error %.3g above 1e-4", err);
        CHECK(gs.coverage == os.coverage, "coverage %lld != %lld", (long long)gs.coverage,
              (long long)os.coverage);
        CHECK(gs.particles == (imap == 0 ? n : n / 2), "particle count %lld",
              (long long)gs.particles);
    }
    BSX(bsx_recon_destroy(h));

    /* extraction from the finished full map: bsx_map_transform + bsx_extract
     * against or_map_transform + or_central_section */
    BSX(bsx_recon_create(&h, &rp));
    BSX(bsx_recon_insert(h, d_imgs, parts, n, NULL));
    BSX(bsx_recon_finish(h, 0, 1, d_map, NULL, NULL, NULL, NULL));
    BSX(bsx_recon_destroy(h));
    HIPC(hipMemcpy(gmap, d_map, sizeof(float) * n3, hipMemcpyDeviceToHost));
    const double origin[3] = {N / 2, N / 2, N / 2};
    float* d_ft;
    HIPC(hipMalloc((void**)&d_ft, sizeof(float) * 2 * n3));
    BSX(bsx_map_transform(d_map, N, origin, d_ft, NULL));
    const int nv = 4, kw = 8, kp = 2;
    double mats[9 * 4];
    for (int v = 0; v < nv; ++v) {
        const double view[4] = {parts[v].view[0], parts[v].view[1], parts[v].view[2],
                                parts[v].view[3]};
        or_view_matrix(view, mats + 9 * v);
    }
    float* d_sec;
    HIPC(hipMalloc((void**)&d_sec, sizeof(float) * 2 * N * N * nv));
    BSX(bsx_extract(d_ft, N, 1.5, mats, nv, 3.0, kw, kp, 1, d_sec, NULL));
    float* gsec = malloc(sizeof(float) * 2 * N * N * nv);
    HIPC(hipMemcpy(gsec, d_sec, sizeof(float) * 2 * N * N * nv, hipMemcpyDeviceToHost));
    double* omap_d = malloc(sizeof(double) * n3);
    for (long i = 0; i < n3; ++i) omap_d[i] = gmap[i];
    double* oft = malloc(sizeof(double) * 2 * n3);
    CHECK(or_map_transform(omap_d, N, origin, oft) == 0, "or_map_transform");
    double* osec = malloc(sizeof(double) * 2 * N * N);
    double worst = 0;
    for (int v = 0; v < nv; ++v) {
        CHECK(or_central_section(oft, N, 1.5, mats + 9 * v, 3.0, kw, kp, 1, osec) == 0,
              "or_central_section");
        const double m = max_abs(osec, 2 * N * N);
        double e = 0;
        for (long i = 0; i < 2 * N * N; ++i)
            e = fmax(e, fabs(gsec[(long)v * 2 * N * N + i] - osec[i]));
        worst = fmax(worst, e / m);
    }
    printf("extraction: %d views, FSI kernel width %d, max rel error %.3g\n", nv, kw, worst);
    CHECK(worst <= 1e-5, "section error %.3g above 1e-5", worst);

    HIPC(hipFree(d_imgs));
    HIPC(hipFree(d_map));
    HIPC(hipFree(d_ft));
    HIPC(hipFree(d_sec));
    free(imgs);
    free(parts);
    free(gmap);
    free(omap);
    free(gsec);
    free(omap_d);
    free(oft);
    free(osec);
    return 0;
}

static int check_errors(void) {
    bsx_recon_params rp;
    memset(&rp, 0, sizeof(rp));
    bsx_recon* h = NULL;
    int r = bsx_recon_create(&h, &rp);            /* size 0 */
    CHECK(r == BSX_ERR_ARG && h == NULL, "create with size 0 returned %d", r);
    CHECK(strlen(bsx_last_error()) > 0, "no error message after a failed create");
    printf("errors: create(size 0) -> %d \"%s\"\n", r, bsx_last_error());
    r = bsx_recon_insert(NULL, NULL, NULL, 1, NULL);
    CHECK(r == BSX_ERR_ARG, "insert on a null handle returned %d", r);
    r = bsx_recon_finish(NULL, 0, 1, NULL, NULL, NULL, NULL, NULL);
    CHECK(r == BSX_ERR_ARG, "finish on a null handle returned %d", r);
    r = bsx_extract(NULL, 32, 1.5, NULL, 1, 3.0, 8, 2, 1, NULL, NULL);
    CHECK(r == BSX_ERR_ARG, "extract with null buffers returned %d", r);
    return 0;
}

int main(void) {
    int major = 0, minor = 0, cus = 0;
    int64_t hbm = 0;
    char name[128];
    if (bsx_version(&major, &minor) != BSX_OK ||
        bsx_device_info(name, sizeof name, &cus, &hbm) != BSX_OK) {
        printf("FAIL device info: %s\n", bsx_last_error());
        return 1;
    }
    printf("libbsx %d.%d on %s, %d CUs, %.0f GB\n", major, minor, name, cus, hbm / 1e9);
    fflush(stdout);
    if (check_errors() || check_reconstruction()) return 1;
    printf(failures ? "bsx_capi_check: %d FAILED\n" : "bsx_capi_check: all checks passed\n",
           failures);
    return failures ? 1 : 0;
}
