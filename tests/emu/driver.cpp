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
// Mode 2 runs the k-mer candidate generation instead (see kmer_mode).
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

// mode 2: k-mer candidates of the set against itself.  h[1] k, h[2] stride,
// h[3] alphabet, h[4] symmetric rule, h[5] ckthr, h[6] encoding; output:
// uint64 count, then the pa_pair candidates.
int kmer_mode(const int32_t *h, const std::vector<uint64_t> &offs,
              const std::vector<uint8_t> &res, const char *outp)
{
    pa_config cfg;
    pa_config_for_cli(&cfg, -11, -2, 6, 0, 2, PA_GAP_AS_CALLED);
    pa_ctx *c = nullptr;
    int rc;
    if ((rc = pa_create(&cfg, &c)))
        return fail("pa_create", rc, nullptr);
    const uint64_t nseq = offs.size() - 1;
    if ((rc = pa_load_seqs(c, PA_SIDE_BOTH, res.data(), offs.data(), nseq,
                           h[6])))
        return fail("pa_load_seqs", rc, c);
    pa_kmer_config kc{};
    kc.k = h[1];
    kc.stride = h[2];
    kc.alphabet = h[3];
    kc.symmetric = h[4];
    kc.ckthr = h[5];
    kc.mosthr = -1.0f;
    uint64_t n = 0;
    if ((rc = pa_kmer_candidates(c, &kc, 0, nseq, nullptr, 0, &n)))
        return fail("pa_kmer_candidates", rc, c);
    void *d = nullptr;
    if ((rc = pa_device_alloc(c, n * sizeof(pa_pair) + 1, &d)))
        return fail("pa_device_alloc", rc, c);
    if ((rc = pa_kmer_candidates(c, &kc, 0, nseq, (pa_pair *)d, n, &n)))
        return fail("pa_kmer_candidates", rc, c);
    std::vector<pa_pair> out(n);
    if (n && (rc = pa_memcpy_d2h(c, out.data(), d, n * sizeof(pa_pair))))
        return fail("pa_memcpy_d2h", rc, c);
    FILE *f = std::fopen(outp, "wb");
    if (!f)
        return 2;
    std::fwrite(&n, sizeof(n), 1, f);
    std::fwrite(out.data(), sizeof(pa_pair), n, f);
    std::fclose(f);
    pa_device_free(c, d);
    pa_destroy(c);
    return 0;
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

    if (h[0] == 2)
        return kmer_mode(h, offs, codes, argv[2]);
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
    // EMU_SHAPES = "s0,s1,s2,s3,e0,e1,e2,stage[,overlap]": row-class kernel
    // shapes, class edges, traceback staging and traceback scheduling
    if (const char *e = std::getenv("EMU_SHAPES")) {
        int v[9];
        const int n = std::sscanf(e, "%d,%d,%d,%d,%d,%d,%d,%d,%d", &v[0],
                                  &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                                  &v[7], &v[8]);
        if (n != 8 && n != 9)
            return 2;
        for (int k = 0; k < 4; ++k)
            t.dp_shape[k] = v[k];
        for (int k = 0; k < 3; ++k)
            t.dp_edge[k] = v[4 + k];
        t.tb_stage = v[7];
        if (n == 9)
            t.tb_overlap = v[8];
    }
    if (const char *e = std::getenv("EMU_XDWAVE"))
        t.xd_wave = std::atoi(e);
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
