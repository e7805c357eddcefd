// pa_io.hip -- the data formats on either side of the aligner: FASTA input
// (pa_fasta_*, pa_load_fasta) and similarity triples / Matrix Market output
// (pa_triples_dev, pa_write_mm).
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pa_ctx.h"

using namespace pa;

// ---------------------------------------------------------------------------
// FASTA normalisation.
//
// One process of the reference reads bytes [0, size-2] of the file
// (src/DistFastaData.cpp:181-188: the last byte is taken to be the final
// newline) and normalises them in place in the FastaData constructor
// (src/FastaData.cpp:13-148).  The observable result, which this parser
// reproduces, is a compacted byte stream S in which
//   - a '>' outside a header line opens a record; the header runs to the
//     next newline;
//   - inside a sequence, a newline is dropped unless the next input byte is
//     '>', '*' is dropped, U/u become T, every other byte is a residue;
//   - a sequence with fewer than k residues is erased from S together with
//     its header and its kept newline, and one more byte: the reference's
//     removal count is one larger than the record (except for a last record
//     that ends in a newline, whose newline is that byte);
// and get_sequence (inc/FastaData.hpp:105-129) reads sequence i as
// S[seq_start_i, id_start_{i+1} - 1), the last one to the end of S, with
// the length taken as a ushort.  So the sequence before an erased record
// that is followed by another record loses its last residue, exactly as in
// the reference.  A first record that is erased reaches one byte before S
// (the reference writes before its buffer); here that byte is absent and
// nothing else changes.
// ---------------------------------------------------------------------------

struct pa_fasta {
    std::vector<char> res;
    std::vector<uint64_t> offs{0};
    std::vector<uint64_t> recs;
    uint64_t wrapped = 0;     // sequences whose length exceeded 65535
};

namespace {

struct FastaParser {
    std::vector<char> S;      // the compacted stream
    std::vector<int64_t> id_start, seq_start;
    std::vector<uint64_t> rec;
    uint64_t nrec = 0;
    // Positions are the reference's buffer indices; S[j] is position
    // j - front (front > 0 once an erased first record reached before 0).
    int64_t front = 0;

    int64_t pos() const { return (int64_t)S.size() - front; }

    // Erase the last record: the reference removes its residues (and the
    // newline ending it, counted as one more residue, when another record
    // follows), its header with the newline after it, and one more byte:
    // normally the byte before its '>'.
    void erase_last_record(int64_t seq_len)
    {
        const int64_t count =
            seq_len + (seq_start.back() - id_start.back()) + 1;
        const int64_t to = pos() - count;   // a position
        if (to + front >= 0) {
            S.resize((size_t)(to + front));
        } else {
            S.clear();
            front = -to;
        }
        id_start.pop_back();
        seq_start.pop_back();
        rec.pop_back();
    }

    void run(const char *b, uint64_t n, int k)
    {
        bool in_name = false, in_seq = false;
        int64_t seq_len = 0;
        S.reserve(n);
        for (uint64_t i = 0; i < n; ++i) {
            char ch = b[i];
            const bool last = i + 1 >= n;
            if (ch == '>' && !in_name) {
                id_start.push_back(pos());
                rec.push_back(nrec++);
                S.push_back(ch);
                seq_len = 0;
                in_name = true;
                in_seq = false;
                continue;
            }
            if (ch == '\n') {
                if (in_name && !last) {
                    S.push_back(ch);
                    seq_start.push_back(pos());
                    in_name = false;
                    in_seq = true;
                } else if (in_seq && !last) {
                    if (b[i + 1] != '>')
                        continue;                // dropped
                    S.push_back(ch);
                    if (seq_len < k)
                        erase_last_record(seq_len + 1);
                } else {
                    S.push_back(ch);
                }
                continue;
            }
            if (ch == '*') {
                if (!in_seq)
                    S.push_back(ch);
                continue;
            }
            if (in_seq) {
                if (ch == 'U' || ch == 'u')
                    ch = 'T';
                ++seq_len;
            }
            S.push_back(ch);
        }
        if (seq_len < k && !id_start.empty() &&
            id_start.size() == seq_start.size())
            erase_last_record(seq_len);
    }
};

}  // namespace

extern "C" {

int pa_fasta_parse(const char *data, uint64_t size, int k, pa_fasta **out)
{
    if (!out || (size && !data) || k < 1)
        return PA_ERR_ARG;
    *out = nullptr;
    pa_fasta *f = new (std::nothrow) pa_fasta();
    if (!f)
        return PA_ERR_NOMEM;
    if (size >= 2) {
        FastaParser p;
        p.run(data, size - 1, k);
        const size_t ns = std::min(p.id_start.size(), p.seq_start.size());
        const int64_t end_all = p.pos();
        for (size_t s = 0; s < ns; ++s) {
            const int64_t end = s + 1 < p.id_start.size() ? p.id_start[s + 1] - 1
                                                          : end_all;
            // a run of erased records can eat past the previous sequence's
            // start (the reference then reads a wrapped length); no residues
            const int64_t span = std::max<int64_t>(0, end - p.seq_start[s]);
            const uint16_t len = (uint16_t)span;   // ushort, as get_sequence
            f->wrapped += span > 65535 ? 1 : 0;
            const int64_t a = p.seq_start[s] + p.front;
            f->res.insert(f->res.end(), p.S.begin() + a, p.S.begin() + a + len);
            f->offs.push_back(f->res.size());
            f->recs.push_back(p.rec[s]);
        }
    }
    *out = f;
    return PA_OK;
}

int pa_fasta_info(const pa_fasta *f, uint64_t *nseq, uint64_t *nres,
                  uint64_t *nwrapped)
{
    if (!f)
        return PA_ERR_ARG;
    if (nseq)
        *nseq = f->recs.size();
    if (nres)
        *nres = f->res.size();
    if (nwrapped)
        *nwrapped = f->wrapped;
    return PA_OK;
}

const char *pa_fasta_residues(const pa_fasta *f)
{
    return f ? f->res.data() : nullptr;
}

const uint64_t *pa_fasta_offsets(const pa_fasta *f)
{
    return f ? f->offs.data() : nullptr;
}

const uint64_t *pa_fasta_records(const pa_fasta *f)
{
    return f ? f->recs.data() : nullptr;
}

void pa_fasta_free(pa_fasta *f) { delete f; }

int pa_load_fasta(pa_ctx *c, int side, const char *data, uint64_t size, int k)
{
    if (!c)
        return PA_ERR_ARG;
    pa_fasta *f = nullptr;
    int rc = pa_fasta_parse(data, size, k, &f);
    if (rc)
        return fail(c, rc, "FASTA parse");
    rc = pa_load_seqs(c, side, (const uint8_t *)f->res.data(), f->offs.data(),
                      f->recs.size(), PA_ENC_ASCII);
    pa_fasta_free(f);
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Similarity triples: the nonzeros that survive the alignment post-prune
// (score_aln != 0, src/DistPWRunner.cpp:149-153) become (row + row_base,
// column + col_base, score_aln), as the drivers collect them for the
// similarity matrix (src/sim_search.cpp:223-232).  Order: the pairs' order.
// ---------------------------------------------------------------------------

namespace {

__global__ void k_triple_flags(const pa_result *__restrict__ res, uint64_t n,
                               uint32_t *__restrict__ flag)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        flag[i] = res[i].score_aln != 0.0f ? 1u : 0u;
}

__global__ void k_triple_write(const pa_pair *__restrict__ pairs,
                               const pa_result *__restrict__ res,
                               const uint32_t *__restrict__ flag,
                               const uint64_t *__restrict__ pos, uint64_t n,
                               uint64_t row_base, uint64_t col_base,
                               pa_triple *__restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (flag[i]) {
            pa_triple t;
            t.row = pairs[i].r + row_base;
            t.col = pairs[i].c + col_base;
            t.score_aln = res[i].score_aln;
            t.pad = 0;
            out[pos[i]] = t;
        }
}

inline unsigned nblk(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}

}  // namespace

extern "C" int pa_triples_dev(pa_ctx *c, const pa_pair *d_pairs,
                              const pa_result *d_res, uint64_t n,
                              uint64_t row_base, uint64_t col_base,
                              pa_triple *d_out, uint64_t *n_out, void *stream)
{
    if (!c || !n_out || (n && (!d_pairs || !d_res || !d_out)))
        return PA_ERR_ARG;
    *n_out = 0;
    if (n == 0)
        return PA_OK;
    hipSetDevice(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->st;
    int rc;
    if ((rc = grow(c, c->tri_flag, n * 4 + 8)) ||
        (rc = grow(c, c->tri_pos, (n + 1) * 8)))
        return rc;
    uint32_t *flag = P<uint32_t>(c->tri_flag);
    uint64_t *pos = P<uint64_t>(c->tri_pos);
    hipLaunchKernelGGL(k_triple_flags, dim3(nblk(n)), dim3(256), 0, s, d_res, n,
                       flag);
    HIPCHK(c, hipGetLastError());
    size_t tmp = 0;
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flag, pos, n + 1,
                                               s));
    if ((rc = grow(c, c->cub, tmp)))
        return rc;
    HIPCHK(c, hipMemsetAsync(flag + n, 0, 4, s));
    HIPCHK(c, hipcub::DeviceScan::ExclusiveSum(c->cub.p, tmp, flag, pos, n + 1,
                                               s));
    hipLaunchKernelGGL(k_triple_write, dim3(nblk(n)), dim3(256), 0, s, d_pairs,
                       d_res, flag, pos, n, row_base, col_base, d_out);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(n_out, pos + n, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    return PA_OK;
}

// ---------------------------------------------------------------------------
// Matrix Market output of the similarity matrix (S.ParallelWriteMM(file,
// true), src/sim_search.cpp:257-269, 626-635; written by CombBLAS, absent
// from the checkout): the coordinate real general format, 1-based indices,
// entries in column-major order, values printed with 6 significant digits
// (the default precision of a C++ stream).  Duplicate coordinates are not
// merged (the caller produces each pair once).
// ---------------------------------------------------------------------------

extern "C" int pa_write_mm(const char *path, uint64_t nrows, uint64_t ncols,
                           const pa_triple *triples, uint64_t n)
{
    if (!path || (n && !triples))
        return PA_ERR_ARG;
    std::vector<uint64_t> ord(n);
    for (uint64_t i = 0; i < n; ++i) {
        if (triples[i].row >= nrows || triples[i].col >= ncols)
            return PA_ERR_INDEX;
        ord[i] = i;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b) {
        const pa_triple &x = triples[a], &y = triples[b];
        return x.col != y.col ? x.col < y.col : x.row < y.row;
    });
    FILE *f = std::fopen(path, "w");
    if (!f)
        return PA_ERR_ARG;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%llu %llu %llu\n", (unsigned long long)nrows,
                 (unsigned long long)ncols, (unsigned long long)n);
    for (uint64_t i : ord)
        std::fprintf(f, "%llu %llu %g\n",
                     (unsigned long long)triples[i].row + 1,
                     (unsigned long long)triples[i].col + 1,
                     (double)triples[i].score_aln);
    const bool ok = std::fflush(f) == 0;
    std::fclose(f);
    return ok ? PA_OK : PA_ERR_ARG;
}
