// bsx_dropin_check.cpp — the drop-in core (bsoft_amd/dropin/bsx_dropin.hpp)
// run the way breconstruct runs particle_reconstruct and
// img_reconstruction_sum_weigh: T = 4 partial maps filled by the reference's
// round-robin, then each output map (nmaps 1, 2 and 3) summed and weighed.
// The weighed transforms and FOM maps are compared with the FP64 oracle
// (oracle/liboracle64.so, the checker) given the same partition.
// Exits 0 only when every check passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../bsoft_amd/dropin/bsx_dropin.hpp"
#include "../../oracle/bsx_oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                                       \
    do {                                                       \
        if (!(cond)) {                                         \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);   \
            std::printf(__VA_ARGS__);                          \
            std::printf("\n");                                 \
            ++failures;                                        \
        }                                                      \
    } while (0)

static uint64_t state = 0xd1a1;
static double urand() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}

int main() {
    const int N = 24, box = 24, n = 90, T = 4;
    bsx_recon_params rp;
    std::memset(&rp, 0, sizeof rp);
    rp.size = N;
    rp.box = box;
    rp.ft_size = bsx_part_ft_size(box, 1.0, 2);
    rp.sampling = 1.5;
    rp.ctf_action = BSX_CTF_WIENER;
    rp.wiener = 0.2;
    rp.nhalves = 1;
    std::vector<float> imgs((size_t)n * box * box);
    std::vector<bsx_particle> parts(n);
    for (int p = 0; p < n; ++p) {
        const double cx = box * (0.35 + 0.3 * urand()), cy = box * (0.35 + 0.3 * urand());
        for (int y = 0; y < box; ++y)
            for (int x = 0; x < box; ++x)
                imgs[((size_t)p * box + y) * box + x] =
                    (float)(urand() - 0.5 + 3 * std::exp(-((x - cx) * (x - cx) + (y - cy) * (y - cy)) / 6.0));
        bsx_particle& q = parts[p];
        std::memset(&q, 0, sizeof q);
        const double a = 2 * M_PI * urand(), z = 2 * urand() - 1, r = std::sqrt(1 - z * z);
        q.view[0] = r * std::cos(a);
        q.view[1] = r * std::sin(a);
        q.view[2] = z;
        q.view[3] = M_PI * (2 * urand() - 1);
        q.ori[0] = q.ori[1] = q.hdr_ori[0] = q.hdr_ori[1] = box / 2;
        q.pixel_size = 1.5;
        q.mag = 1;
        q.volt = 300e3;
        q.cs = 2.7e7;
        q.amp_shift = std::asin(0.07);
        q.def_avg = 1.2e4 + 1.5e4 * urand();
        q.sel = 1;
    }
    // round-robin: the k-th selected particle goes to partial map (k-1) mod T
    std::vector<std::vector<int>> lists(T);
    for (int k = 1; k <= n; ++k) lists[bsx_dropin::partial_index(k, 1, T)].push_back(k - 1);
    const size_t n3 = (size_t)N * N * N;
    std::vector<std::vector<float>> store(T, std::vector<float>(5 * n3));
    std::vector<bsx_dropin::PartialPlanes> pacc(T);
    for (int t = 0; t < T; ++t) {
        bsx_dropin::PartialMap pm(rp, 16);
        std::vector<float> li;
        std::vector<bsx_particle> lp;
        for (int g : lists[t]) {
            li.insert(li.end(), imgs.begin() + (size_t)g * box * box,
                      imgs.begin() + (size_t)(g + 1) * box * box);
            lp.push_back(parts[g]);
        }
        float* s = store[t].data();
        int r = pm.insert(li.data(), lp.data(), (long)lp.size());
        if (!r) r = pm.planes(s, s + 2 * n3, s + 3 * n3, s + 4 * n3);
        if (r) {
            std::printf("FAIL partial map %d: %d %s\n", t, r, pm.message().c_str());
            return 1;
        }
        pacc[t] = {s, s + 2 * n3, s + 3 * n3, s + 4 * n3, pm.count()};
    }
    // the oracle with the same T partial maps (its own round-robin)
    or_recon* o = or_recon_new(&rp, T);
    CHECK(o && or_recon_insert(o, parts.data(), imgs.data(), n, 1) == 0, "oracle insert");
    std::vector<float> ft(2 * n3), fom(n3);
    std::vector<double> oft(2 * n3), ofom(n3);
    const int cases[][2] = {{0, 1}, {0, 2}, {1, 2}, {0, 3}, {1, 3}, {2, 3}};
    for (auto& c : cases) {
        const int imap = c[0], nmaps = c[1];
        bsx_stats st;
        std::string msg;
        int r = bsx_dropin::sum_weigh(rp, pacc.data(), imap, nmaps, T, ft.data(), fom.data(), &st,
                                      &msg);
        if (r) {
            std::printf("FAIL sum_weigh(%d, %d): %d %s\n", imap, nmaps, r, msg.c_str());
            return 1;
        }
        bsx_stats os;
        CHECK(or_recon_finish(o, imap, nmaps, nullptr, ofom.data(), oft.data(), &os) == 0,
              "oracle finish");
        double m = 0, e = 0;
        for (size_t i = 0; i < 2 * n3; ++i) m = std::fmax(m, std::fabs(oft[i]));
        for (size_t i = 0; i < 2 * n3; ++i) e = std::fmax(e, std::fabs(ft[i] - oft[i]));
        // FOM: compare where the oracle's value is well conditioned
        double fe = 0;
        long nf = 0;
        for (size_t i = 0; i < n3; ++i)
            if (std::fabs(ofom[i]) > 1e-2 && std::fabs(ofom[i]) < 1e3) {
                fe = std::fmax(fe, std::fabs(fom[i] - ofom[i]) / std::fabs(ofom[i]));
                ++nf;
            }
        std::printf("sum_weigh imap %d nmaps %d: %lld particles, coverage %lld (oracle %lld), "
                    "transform max rel error %.3g, FOM max rel error %.3g over %ld voxels\n",
                    imap, nmaps, (long long)st.particles, (long long)st.coverage,
                    (long long)os.coverage, e / m, fe, nf);
        CHECK(e / m <= 1e-4, "transform error %.3g", e / m);
        CHECK(st.coverage == os.coverage, "coverage");
        CHECK(fe <= 2e-3, "FOM error %.3g", fe);
    }
    or_recon_free(o);
    // half sets follow the partial maps: T even, half = partial % 2
    CHECK(bsx_dropin::half_of(bsx_dropin::partial_index(1, 1, T), T) == 0 &&
              bsx_dropin::half_of(bsx_dropin::partial_index(2, 1, T), T) == 1 &&
              bsx_dropin::half_of(bsx_dropin::partial_index(5, 1, T), T) == 0,
          "half assignment");
    std::printf(failures ? "bsx_dropin_check: %d FAILED\n" : "bsx_dropin_check: all checks passed\n",
                failures);
    return failures ? 1 : 0;
}
