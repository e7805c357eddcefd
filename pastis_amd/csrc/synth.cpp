// synth.cpp -- deterministic synthetic protein sets and candidate pairs.
//
// Builds the inputs of the benchmark configurations: proteins drawn from the
// Robinson & Robinson background frequencies, homolog families made by
// mutating a root under BLOSUM62 substitution odds P(b|a) ~ p_b 2^(S(a,b)/2)
// with geometric indels, and candidate pairs carrying the k-mer fields of
// CommonKmerLoc (count, max k-mer self-score, first/second seed offsets) as
// the KmerIntersectLoc semiring defines them (inc/sr.hpp:54-78): count =
// number of shared k-mers, first = offsets of the first shared k-mer (lowest
// k-mer code), second = offsets of the last one, score = max self-score.
// Pairs that share no exact k-mer get one planted seed (count 1, second =
// (0,0) as the default-constructed pair of inc/Types.hpp:170-171): on the
// true alignment for homologs, at random offsets otherwise.
//
// Everything uses std::mt19937_64 with explicit seeds.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

const char kAlphabet[] = "ARNDCQEGHILKMFPSTWYVBZX*J";
// Robinson & Robinson (1991), order ARNDCQEGHILKMFPSTWYV
const double kBg[20] = {0.07805, 0.05129, 0.04487, 0.05364, 0.01925,
                        0.04264, 0.06295, 0.07377, 0.02199, 0.05142,
                        0.09019, 0.05744, 0.02243, 0.03856, 0.05203,
                        0.07120, 0.05841, 0.01330, 0.03216, 0.06441};
// BLOSUM62 over the 20 standard residues (inc/ScoreMat.hpp:104-131)
const int8_t kB62[20][20] = {
    {4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0},
    {-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3},
    {-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3},
    {-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3},
    {0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1},
    {-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2},
    {-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2},
    {0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3},
    {-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3},
    {-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3},
    {-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1},
    {-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2},
    {-1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1},
    {-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1},
    {-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2},
    {1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2},
    {0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0},
    {-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3},
    {-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1},
    {0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4}};

struct Pair {  // pa_pair layout
    uint32_t r, c;
    uint16_t count;
    int16_t kscore;
    uint16_t seed[4];
};
static_assert(sizeof(Pair) == 20, "pair layout");

struct Set {
    std::vector<uint8_t> codes;
    std::vector<uint64_t> offs;
    std::vector<int32_t> family;          // family id per sequence
    std::vector<int32_t> root;            // root sequence of the family
    std::vector<std::vector<int32_t>> to_root;  // residue -> root position
    std::vector<std::vector<int32_t>> members;  // per family
    std::vector<Pair> pairs;
    double mean_identity = 0;
};

struct Sampler {
    std::mt19937_64 rng;
    std::discrete_distribution<int> bg;
    std::vector<std::discrete_distribution<int>> sub;
    explicit Sampler(uint64_t seed) : rng(seed), bg(kBg, kBg + 20)
    {
        for (int a = 0; a < 20; ++a) {
            std::vector<double> w(20);
            for (int b = 0; b < 20; ++b)
                w[b] = b == a ? 0.0 : kBg[b] * std::pow(2.0, kB62[a][b] / 2.0);
            sub.emplace_back(w.begin(), w.end());
        }
    }
    double uni() { return std::uniform_real_distribution<double>(0, 1)(rng); }
    int uint(int lo, int hi)
    {
        return std::uniform_int_distribution<int>(lo, hi)(rng);
    }
};

int draw_len(Sampler &S, int model, double a, double b, double tail_frac,
             double tail_lo, double tail_hi)
{
    if (tail_frac > 0 && S.uni() < tail_frac)
        return (int)std::floor(tail_lo + S.uni() * (tail_hi - tail_lo + 1));
    if (model == 0)  // uniform [a, b]
        return S.uint((int)a, (int)b);
    // log-normal with median a and sigma b, clipped to [30, 3000] by caller
    std::lognormal_distribution<double> ln(std::log(a), b);
    return (int)std::lround(ln(S.rng));
}

// Mutates root into a child of roughly `ident` identity; map gives for each
// child residue the root position it descends from (-1 for insertions).
void mutate(Sampler &S, const uint8_t *root, int n, double ident,
            double indel_p, std::vector<uint8_t> &out,
            std::vector<int32_t> &map)
{
    out.clear();
    map.clear();
    const double ev = indel_p * (1.0 - ident);
    // per-site substitution probability giving the target identity; a drawn
    // substitute is always a different residue
    const double sub_p = std::min(1.0, std::max(0.0, 1.0 - ident));
    for (int i = 0; i < n; ++i) {
        if (S.uni() < ev) {
            int L = 1;
            while (S.uni() < 0.5 && L < 30)
                ++L;
            if (S.uni() < 0.5) {  // insertion before position i
                for (int k = 0; k < L; ++k) {
                    out.push_back((uint8_t)S.bg(S.rng));
                    map.push_back(-1);
                }
            } else {  // deletion of L root residues
                i += L - 1;
                continue;
            }
        }
        uint8_t a = root[i];
        if (a < 20 && S.uni() < sub_p)
            a = (uint8_t)S.sub[a](S.rng);
        out.push_back(a);
        map.push_back(i);
    }
    if (out.empty()) {
        out.push_back(root[0]);
        map.push_back(0);
    }
}

uint64_t kcode(const uint8_t *s, int k)
{
    uint64_t c = 0;
    for (int i = 0; i < k; ++i)
        c = c * 25 + s[i];
    return c;
}

int self_score(const uint8_t *s, int k)
{
    int v = 0;
    for (int i = 0; i < k; ++i)
        v += s[i] < 20 ? kB62[s[i]][s[i]] : 0;
    return v;
}

struct KList {  // sorted (code, first offset) of one sequence
    std::vector<std::pair<uint64_t, uint32_t>> v;
};

void kmers(const uint8_t *s, int n, int k, KList &out)
{
    out.v.clear();
    for (int i = 0; i + k <= n; ++i)
        out.v.push_back({kcode(s + i, k), (uint32_t)i});
    std::stable_sort(out.v.begin(), out.v.end(),
                     [](const std::pair<uint64_t, uint32_t> &a,
                        const std::pair<uint64_t, uint32_t> &b) {
                         return a.first < b.first;
                     });
    // keep the first occurrence of each code
    out.v.erase(std::unique(out.v.begin(), out.v.end(),
                            [](const std::pair<uint64_t, uint32_t> &a,
                               const std::pair<uint64_t, uint32_t> &b) {
                                return a.first == b.first;
                            }),
                out.v.end());
}

}  // namespace

extern "C" {

// model: 0 = uniform [a, b], 1 = log-normal (median a, sigma b) clipped to
// [clip_lo, clip_hi].  fam_size members per family (1 = no homologs), the
// first of each family is its root; children get identity U[id_lo, id_hi].
void *sg_new_set(uint64_t nseq, int model, double a, double b, int clip_lo,
                 int clip_hi, double tail_frac, double tail_lo, double tail_hi,
                 int fam_size, double id_lo, double id_hi, double indel_p,
                 uint64_t seed)
{
    Set *S = new Set();
    Sampler sm(seed);
    S->offs.push_back(0);
    S->family.resize(nseq);
    S->root.resize(nseq);
    S->to_root.resize(nseq);
    std::vector<uint8_t> child;
    std::vector<int32_t> map;
    double id_sum = 0;
    uint64_t id_n = 0;
    for (uint64_t i = 0; i < nseq; ++i) {
        const int fam = (int)(i / (uint64_t)std::max(fam_size, 1));
        const uint64_t r = (uint64_t)fam * (uint64_t)std::max(fam_size, 1);
        S->family[i] = fam;
        S->root[i] = (int32_t)r;
        if ((int)S->members.size() <= fam)
            S->members.resize(fam + 1);
        S->members[fam].push_back((int32_t)i);
        if (i == r) {
            int L = draw_len(sm, model, a, b, tail_frac, tail_lo, tail_hi);
            if (!(tail_frac > 0 && L >= tail_lo))
                L = std::min(std::max(L, clip_lo), clip_hi);
            for (int k = 0; k < L; ++k)
                S->codes.push_back((uint8_t)sm.bg(sm.rng));
            S->to_root[i].resize(L);
            for (int k = 0; k < L; ++k)
                S->to_root[i][k] = k;
        } else {
            const double ident = id_lo + sm.uni() * (id_hi - id_lo);
            id_sum += ident;
            ++id_n;
            const uint8_t *rp = S->codes.data() + S->offs[r];
            const int rn = (int)(S->offs[r + 1] - S->offs[r]);
            mutate(sm, rp, rn, ident, indel_p, child, map);
            if (child.size() > 65535) {
                child.resize(65535);
                map.resize(65535);
            }
            S->codes.insert(S->codes.end(), child.begin(), child.end());
            S->to_root[i] = map;
        }
        S->offs.push_back(S->codes.size());
    }
    S->mean_identity = id_n ? id_sum / id_n : 0;
    return S;
}

// Families of random sizes: sequences are drawn log-normal (median a,
// sigma b, clipped to [clip_lo, clip_hi]); families of U[fam_lo, fam_hi]
// members (a root and mutated children at identity U[id_lo, id_hi]) are
// formed until they hold fam_frac of the nseq sequences, the rest are
// unrelated roots.  Families come first, in order.
void *sg_new_families(uint64_t nseq, double a, double b, int clip_lo,
                      int clip_hi, double fam_frac, int fam_lo, int fam_hi,
                      double id_lo, double id_hi, double indel_p,
                      uint64_t seed)
{
    Set *S = new Set();
    Sampler sm(seed);
    S->offs.push_back(0);
    S->family.resize(nseq);
    S->root.resize(nseq);
    S->to_root.resize(nseq);
    std::vector<uint8_t> child;
    std::vector<int32_t> map;
    double id_sum = 0;
    uint64_t id_n = 0;
    const uint64_t in_fam = (uint64_t)std::llround(fam_frac * (double)nseq);
    uint64_t i = 0;
    int fam = 0;
    auto add_root = [&](uint64_t at) {
        int L = draw_len(sm, 1, a, b, 0.0, 0.0, 0.0);
        L = std::min(std::max(L, clip_lo), clip_hi);
        for (int k = 0; k < L; ++k)
            S->codes.push_back((uint8_t)sm.bg(sm.rng));
        S->to_root[at].resize(L);
        for (int k = 0; k < L; ++k)
            S->to_root[at][k] = k;
        S->offs.push_back(S->codes.size());
    };
    while (i < nseq) {
        uint64_t size = 1;
        if (i < in_fam) {
            size = (uint64_t)sm.uint(fam_lo, fam_hi);
            size = std::min<uint64_t>(size, in_fam - i);
            if (size < 2 && i + 1 < in_fam)
                size = 2;
        }
        const uint64_t r = i;
        if ((int)S->members.size() <= fam)
            S->members.resize(fam + 1);
        for (uint64_t m = 0; m < size && i < nseq; ++m, ++i) {
            S->family[i] = fam;
            S->root[i] = (int32_t)r;
            S->members[fam].push_back((int32_t)i);
            if (m == 0) {
                add_root(i);
                continue;
            }
            const double ident = id_lo + sm.uni() * (id_hi - id_lo);
            id_sum += ident;
            ++id_n;
            const uint8_t *rp = S->codes.data() + S->offs[r];
            const int rn = (int)(S->offs[r + 1] - S->offs[r]);
            mutate(sm, rp, rn, ident, indel_p, child, map);
            if (child.size() > 65535) {
                child.resize(65535);
                map.resize(65535);
            }
            S->codes.insert(S->codes.end(), child.begin(), child.end());
            S->to_root[i] = map;
            S->offs.push_back(S->codes.size());
        }
        ++fam;
    }
    S->mean_identity = id_n ? id_sum / id_n : 0;
    return S;
}

// Candidate pairs: frac_hom of them within a family (r != c), the rest
// uniform (zipf_s <= 0) or with both ends drawn from a Zipf(zipf_s) law over a
// random permutation of the sequences; r != c always.
void sg_make_pairs(void *h, uint64_t npairs, double frac_hom, double zipf_s,
                   int k, int nthreads, uint64_t seed)
{
    Set *S = (Set *)h;
    const uint64_t nseq = S->offs.size() - 1;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0, 1);
    std::vector<uint32_t> perm(nseq);
    for (uint64_t i = 0; i < nseq; ++i)
        perm[i] = (uint32_t)i;
    std::vector<double> cdf;
    if (zipf_s > 0) {
        std::shuffle(perm.begin(), perm.end(), rng);
        cdf.resize(nseq);
        double acc = 0;
        for (uint64_t i = 0; i < nseq; ++i) {
            acc += 1.0 / std::pow((double)(i + 1), zipf_s);
            cdf[i] = acc;
        }
        for (auto &x : cdf)
            x /= acc;
    }
    auto draw = [&]() -> uint32_t {
        if (zipf_s > 0) {
            const double u = U(rng);
            uint64_t i = std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
            if (i >= nseq)
                i = nseq - 1;
            return perm[i];
        }
        return (uint32_t)(rng() % nseq);
    };
    S->pairs.resize(npairs);
    std::vector<uint8_t> hom(npairs);
    for (uint64_t t = 0; t < npairs; ++t) {
        Pair &p = S->pairs[t];
        std::memset(&p, 0, sizeof(p));
        const bool is_hom = U(rng) < frac_hom;
        hom[t] = 0;
        if (is_hom) {
            const uint32_t a = draw();
            const auto &mem = S->members[S->family[a]];
            if (mem.size() > 1) {
                uint32_t b = a;
                while (b == a)
                    b = (uint32_t)mem[rng() % mem.size()];
                p.r = a;
                p.c = b;
                hom[t] = 1;
                continue;
            }
        }
        uint32_t a = draw(), b = draw();
        while (b == a)
            b = draw();
        p.r = a;
        p.c = b;
    }
    // k-mer fields
    std::vector<uint64_t> pseed(npairs);
    for (uint64_t t = 0; t < npairs; ++t)
        pseed[t] = rng();
    std::vector<KList> kl(nseq);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long i = 0; i < (long long)nseq; ++i)
        kmers(S->codes.data() + S->offs[i], (int)(S->offs[i + 1] - S->offs[i]),
              k, kl[i]);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long t = 0; t < (long long)npairs; ++t) {
        Pair &p = S->pairs[t];
        const KList &A = kl[p.r], &B = kl[p.c];
        const uint8_t *sr = S->codes.data() + S->offs[p.r];
        const int nr = (int)(S->offs[p.r + 1] - S->offs[p.r]);
        const int nc = (int)(S->offs[p.c + 1] - S->offs[p.c]);
        size_t i = 0, j = 0;
        uint32_t cnt = 0;
        int best = -1;
        while (i < A.v.size() && j < B.v.size()) {
            if (A.v[i].first < B.v[j].first) {
                ++i;
            } else if (A.v[i].first > B.v[j].first) {
                ++j;
            } else {
                const uint16_t ro = (uint16_t)A.v[i].second;
                const uint16_t co = (uint16_t)B.v[j].second;
                if (cnt == 0) {
                    p.seed[0] = ro;
                    p.seed[1] = co;
                } else {
                    p.seed[2] = ro;
                    p.seed[3] = co;
                }
                best = std::max(best, self_score(sr + ro, k));
                ++cnt;
                ++i;
                ++j;
            }
        }
        if (cnt > 0) {
            p.count = (uint16_t)std::min<uint32_t>(cnt, 65535);
            p.kscore = (int16_t)best;
            continue;
        }
        // planted seed
        std::mt19937_64 lr(pseed[t]);
        int ro = -1, co = -1;
        if (hom[t]) {
            // a residue pair descending from the same root position
            const auto &mr = S->to_root[p.r];
            const auto &mc = S->to_root[p.c];
            std::vector<int32_t> inv(mc.size() ? 1 + *std::max_element(mc.begin(), mc.end()) : 0, -1);
            for (int x = 0; x < (int)mc.size(); ++x)
                if (mc[x] >= 0 && mc[x] < (int)inv.size() && inv[mc[x]] < 0)
                    inv[mc[x]] = x;
            std::vector<std::pair<int, int>> cand;
            for (int x = 0; x < (int)mr.size(); ++x)
                if (mr[x] >= 0 && mr[x] < (int)inv.size() && inv[mr[x]] >= 0 &&
                    x + k <= nr && inv[mr[x]] + k <= nc)
                    cand.push_back({x, inv[mr[x]]});
            if (!cand.empty()) {
                auto pr = cand[lr() % cand.size()];
                ro = pr.first;
                co = pr.second;
            }
        }
        if (ro < 0) {
            ro = nr >= k ? (int)(lr() % (uint64_t)(nr - k + 1)) : 0;
            co = nc >= k ? (int)(lr() % (uint64_t)(nc - k + 1)) : 0;
        }
        p.count = 1;
        p.seed[0] = (uint16_t)ro;
        p.seed[1] = (uint16_t)co;
        p.seed[2] = p.seed[3] = 0;
        p.kscore = (int16_t)self_score(sr + ro, k);
    }
}

uint64_t sg_num_seqs(void *h) { return ((Set *)h)->offs.size() - 1; }
uint64_t sg_total_residues(void *h) { return ((Set *)h)->codes.size(); }
uint64_t sg_num_pairs(void *h) { return ((Set *)h)->pairs.size(); }
double sg_mean_identity(void *h) { return ((Set *)h)->mean_identity; }

void sg_copy_seqs(void *h, uint8_t *codes, uint64_t *offs)
{
    Set *S = (Set *)h;
    std::memcpy(codes, S->codes.data(), S->codes.size());
    std::memcpy(offs, S->offs.data(), S->offs.size() * sizeof(uint64_t));
}

void sg_copy_pairs(void *h, void *out)
{
    Set *S = (Set *)h;
    std::memcpy(out, S->pairs.data(), S->pairs.size() * sizeof(Pair));
}

void sg_copy_family(void *h, int32_t *fam)
{
    Set *S = (Set *)h;
    std::memcpy(fam, S->family.data(), S->family.size() * sizeof(int32_t));
}

void sg_free(void *h) { delete (Set *)h; }

}  // extern "C"
