// driver.cpp -- TEST ONLY.  Runs one batch through the C ABI of the CPU
// emulation build of libpastis_hip (tests/emu), so that the library's own
// kernels and host code run on the CPU, also under AddressSanitizer and
// UndefinedBehaviorSanitizer.
//
//   driver <input> <output>
// input : int32 mode, gap_open_cli, gap_ext_cli, klength, xdrop, seed_count,
//         gap_model, xd_coop; uint64 nseq, npairs; uint64 offsets[nseq + 1];
//         uint8 codes[offsets[nseq]]; pa_pair pairs[npairs]
// output: pa_result results[npairs], then (X-drop mode) pa_extent of seed 0
//         and of seed 1 for every pair, then the uint64 X-drop cell count.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/pastis_hip.h"

namespace {

template <class T> bool rd(FILE *f, T *p, size_t n)
{
    return std::fread(p, sizeof(T), n, f) == n;
}

int fail(const char *what, int rc, pa_ctx *c)
{
    std::fprintf(stderr, "driver: %s: %s (%s)\n", what, pa_strerror(rc),
                 c ? pa_last_error(c) : "");
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: driver <input> <output>\n");
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in)
        return 2;
    int32_t h[8];
    uint64_t nseq = 0, np = 0;
    if (!rd(in, h, 8) || !rd(in, &nseq, 1) || !rd(in, &np, 1))
        return 2;
    std::vector<uint64_t> offs(nseq + 1);
    if (!rd(in, offs.data(), offs.size()))
        return 2;
    std::vector<uint8_t> codes(offs[nseq]);
    std::vector<pa_pair> pairs(np);
    if (!rd(in, codes.data(), codes.size()) || !rd(in, pairs.data(), np))
        return 2;
    std::fclose(in);

    pa_config cfg;
    int rc = pa_config_for_cli(&cfg, h[1], h[2], h[3], h[4], h[5], h[6]);
    if (rc)
        return fail("pa_config_for_cli", rc, nullptr);
    cfg.mode = h[0];
    pa_ctx *c = nullptr;
    if ((rc = pa_create(&cfg, &c)))
        return fail("pa_create", rc, nullptr);
    pa_tuning t;
    pa_get_tuning(c, &t);
    t.xd_coop = h[7];
    if ((rc = pa_set_tuning(c, &t)))
        return fail("pa_set_tuning", rc, c);
    if ((rc = pa_load_seqs(c, PA_SIDE_BOTH, codes.data(), offs.data(), nseq,
                           PA_ENC_CODES)))
        return fail("pa_load_seqs", rc, c);
    std::vector<pa_result> res(np);
    if ((rc = pa_align(c, pairs.data(), np, res.data())))
        return fail("pa_align", rc, c);
    FILE *out = std::fopen(argv[2], "wb");
    if (!out)
        return 2;
    std::fwrite(res.data(), sizeof(pa_result), np, out);
    if (cfg.mode == PA_MODE_XDROP) {
        std::vector<pa_extent> ext(np);
        for (int s = 0; s < cfg.seed_cnt; ++s) {
            if ((rc = pa_xdrop_extents(c, pairs.data(), np, s, ext.data())))
                return fail("pa_xdrop_extents", rc, c);
            std::fwrite(ext.data(), sizeof(pa_extent), np, out);
        }
        pa_stats st;
        pa_get_stats(c, &st);
        std::fwrite(&st.xdrop_cells, sizeof(uint64_t), 1, out);
    }
    std::fclose(out);
    pa_destroy(c);
    return 0;
}
