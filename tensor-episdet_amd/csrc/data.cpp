// data.cpp -- dataset format (reference .txt + .bin files) and the seeded
// synthetic generator.  Host code only; built with g++ into libepisdet_data.so
// and also linked into the CLI.
//
// Format follows the reference's loader, tensor-episdet.cu:58-148 (five-line
// .txt, controls first, .bin paths relative to the .txt, raw u32 planes padded
// to a multiple of 1024 samples, MSB-first bit order).
#include "episdet_data.h"

#include <libgen.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

inline uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** seeded from (seed, stream, sub) through splitmix64.
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream, uint64_t sub)
    {
        uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ (sub * 0x8CB92BA72F3D8DD7ull);
        for (int i = 0; i < 4; i++) s[i] = splitmix(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    inline uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Marsaglia-Tsang gamma(shape, 1); shape < 1 via the U^(1/shape) boost.
    double gamma(double shape)
    {
        if (shape < 1.0) {
            double u = uniform();
            return gamma(shape + 1.0) * std::pow(u > 0 ? u : 1e-300, 1.0 / shape);
        }
        double d = shape - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                // Box-Muller normal
                double u1 = uniform(), u2 = uniform();
                if (u1 <= 0) u1 = 1e-300;
                x = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
                v = 1.0 + c * x;
            } while (v <= 0);
            v = v * v * v;
            double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
            if (std::log(u > 0 ? u : 1e-300) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
        }
    }
};

struct Hwe {
    uint32_t t0, t1;  // u < t0 -> genotype 0; u < t1 -> 1; else 2
    explicit Hwe(double p)
    {
        double q = 1.0 - p;
        double a = q * q, b = a + 2.0 * p * q;
        t0 = (uint32_t)std::min(4294967295.0, std::floor(a * 4294967296.0));
        t1 = (uint32_t)std::min(4294967295.0, std::floor(b * 4294967296.0));
    }
    inline int draw(uint32_t u) const { return u < t0 ? 0 : (u < t1 ? 1 : 2); }
};

inline void set_geno(uint32_t *row0, uint32_t *row1, int s, int g)
{
    uint32_t bit = 1u << (31 - (s & 31));
    if (g == 0) row0[s >> 5] |= bit;
    else if (g == 1) row1[s >> 5] |= bit;
}

}  // namespace

extern "C" {

int episdet_padded_words(int num_samples) { return 32 * ((num_samples + 1023) / 1024); }

void episdet_gen_defaults(episdet_gen_params *gp)
{
    std::memset(gp, 0, sizeof *gp);
    gp->seed = 1;
    gp->maf_dist = EPISDET_MAF_UNIFORM;
    gp->maf_lo = 0.05;
    gp->maf_hi = 0.5;
    gp->beta_a = 0.5;
    gp->beta_b = 2.0;
    gp->planted_sets = 0;
    gp->planted_order = 3;
    gp->baseline_risk = 0.1;
    gp->peak_risk = 0.4;
    gp->planted_maf_lo = 0.25;
    gp->planted_maf_hi = 0.5;
}

int episdet_generate(const episdet_gen_params *gp, uint32_t *cases, uint32_t *controls, int *planted,
                     double *penetrance)
{
    if (!gp || !cases || !controls) return 1;
    const int n = gp->num_snps, Nc = gp->num_cases, Nt = gp->num_controls;
    const int Wc = gp->cases_words ? gp->cases_words : episdet_padded_words(Nc);
    const int Wt = gp->controls_words ? gp->controls_words : episdet_padded_words(Nt);
    const int order = gp->planted_order;
    const int nsets = gp->planted_sets;
    if (n <= 0 || Nc <= 0 || Nt <= 0 || Wc * 32 < Nc || Wt * 32 < Nt) return 1;
    if (nsets < 0 || nsets > 2 || (nsets > 0 && (order < 2 || order > 3 || n < nsets * order))) return 1;

    // Per-SNP minor allele frequencies (stream 0).
    std::vector<double> maf(n);
    {
        Rng r(gp->seed, 0, 0);
        for (int v = 0; v < n; v++) {
            double p;
            if (gp->maf_dist == EPISDET_MAF_BETA) {
                double a = r.gamma(gp->beta_a), b = r.gamma(gp->beta_b);
                p = a / (a + b);
                p = std::min(std::max(p, gp->maf_lo), gp->maf_hi);
            } else {
                p = gp->maf_lo + (gp->maf_hi - gp->maf_lo) * r.uniform();
            }
            maf[v] = p;
        }
    }

    // Planted sets (stream 1): distinct SNPs, sorted within each set.
    std::vector<int> psnp;  // nsets*order
    std::vector<double> pen((size_t)std::max(nsets, 1) * 27, gp->baseline_risk);
    if (nsets > 0) {
        Rng r(gp->seed, 1, 0);
        std::vector<char> used(n, 0);
        for (int s = 0; s < nsets; s++) {
            std::vector<int> set;
            while ((int)set.size() < order) {
                int v = (int)(r.next() % (uint64_t)n);
                if (used[v]) continue;
                used[v] = 1;
                set.push_back(v);
            }
            std::sort(set.begin(), set.end());
            for (int v : set) {
                psnp.push_back(v);
                maf[v] = gp->planted_maf_lo + (gp->planted_maf_hi - gp->planted_maf_lo) * r.uniform();
            }
        }
        // Penetrance (stream 2): cells where every planted SNP carries a minor
        // allele get baseline + (peak - baseline) * U[0.5, 1]; the all-homozygous
        // minor cell gets the peak exactly; every other cell the baseline.
        Rng rp(gp->seed, 2, 0);
        int ncell = order == 3 ? 27 : 9;
        for (int s = 0; s < nsets; s++) {
            for (int c = 0; c < ncell; c++) {
                int g[3] = {c / 9, (c / 3) % 3, c % 3};
                if (order == 2) { g[0] = c / 3; g[1] = c % 3; g[2] = 1; }
                bool risk = g[0] >= 1 && g[1] >= 1 && g[2] >= 1;
                double u = rp.uniform();
                pen[(size_t)s * 27 + c] =
                    risk ? gp->baseline_risk + (gp->peak_risk - gp->baseline_risk) * (0.5 + 0.5 * u)
                         : gp->baseline_risk;
            }
            pen[(size_t)s * 27 + (ncell - 1)] = gp->peak_risk;
        }
    }
    if (planted) {
        for (int s = 0; s < nsets; s++)
            for (int o = 0; o < 3; o++) planted[s * 3 + o] = o < order ? psnp[(size_t)s * order + o] : -1;
    }
    if (penetrance) {
        for (int s = 0; s < nsets; s++)
            for (int c = 0; c < 27; c++) penetrance[s * 27 + c] = pen[(size_t)s * 27 + c];
    }

    std::memset(cases, 0, sizeof(uint32_t) * (size_t)n * 2 * Wc);
    std::memset(controls, 0, sizeof(uint32_t) * (size_t)n * 2 * Wt);

    // Status sampling for planted sets (stream 3): planted genotypes per kept
    // individual, per class.
    const int np = (int)psnp.size();
    std::vector<uint8_t> pg_cas((size_t)np * Nc), pg_ctl((size_t)np * Nt);
    if (np > 0) {
        Rng r(gp->seed, 3, 0);
        std::vector<Hwe> h;
        for (int v : psnp) h.emplace_back(maf[v]);
        std::vector<int> g(np);
        int kc = 0, kt = 0;
        while (kc < Nc || kt < Nt) {
            for (int i = 0; i < np; i++) g[i] = h[i].draw((uint32_t)(r.next() >> 32));
            double risk = gp->baseline_risk;
            for (int s = 0; s < nsets; s++) {
                int c = 0;
                for (int o = 0; o < order; o++) c = c * 3 + g[s * order + o];
                risk += pen[(size_t)s * 27 + c] - gp->baseline_risk;
            }
            risk = std::min(1.0, std::max(0.0, risk));
            bool is_case = r.uniform() < risk;
            if (is_case && kc < Nc) {
                for (int i = 0; i < np; i++) pg_cas[(size_t)i * Nc + kc] = (uint8_t)g[i];
                kc++;
            } else if (!is_case && kt < Nt) {
                for (int i = 0; i < np; i++) pg_ctl[(size_t)i * Nt + kt] = (uint8_t)g[i];
                kt++;
            }
        }
    }
    std::vector<int> pidx(n, -1);
    for (int i = 0; i < np; i++) pidx[psnp[i]] = i;

    if (gp->threads > 0) omp_set_num_threads(gp->threads);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int v = 0; v < n; v++) {
        for (int cls = 0; cls < 2; cls++) {
            const int N = cls == 0 ? Nc : Nt, W = cls == 0 ? Wc : Wt;
            uint32_t *row0 = (cls == 0 ? cases : controls) + (size_t)v * 2 * W;
            uint32_t *row1 = row0 + W;
            if (pidx[v] >= 0) {
                const uint8_t *pg = (cls == 0 ? pg_cas.data() + (size_t)pidx[v] * Nc
                                              : pg_ctl.data() + (size_t)pidx[v] * Nt);
                for (int s = 0; s < N; s++) set_geno(row0, row1, s, pg[s]);
                continue;
            }
            // Stream 4: independent HWE genotypes, one stream per (SNP, class).
            Rng r(gp->seed, 4, (uint64_t)v * 2 + cls);
            Hwe h(maf[v]);
            for (int s = 0; s < N; s += 2) {
                uint64_t u = r.next();
                set_geno(row0, row1, s, h.draw((uint32_t)(u >> 32)));
                if (s + 1 < N) set_geno(row0, row1, s + 1, h.draw((uint32_t)u));
            }
        }
    }
    return 0;
}

int episdet_write_dataset(const char *dir, const char *name, int num_snps, int num_cases, int num_controls,
                          const uint32_t *cases, const uint32_t *controls)
{
    std::string d(dir), nm(name);
    std::string txt = d + "/" + nm + ".txt";
    std::string fc = nm + ".cases.bin", ft = nm + ".controls.bin";
    FILE *f = std::fopen(txt.c_str(), "w");
    if (!f) return 1;
    std::fprintf(f, "%d\n%s\n%d\n%s\n%d\n", num_snps, ft.c_str(), num_controls, fc.c_str(), num_cases);
    std::fclose(f);
    const int Wc = episdet_padded_words(num_cases), Wt = episdet_padded_words(num_controls);
    struct {
        const std::string &file;
        const uint32_t *p;
        int W;
    } out[2] = {{fc, cases, Wc}, {ft, controls, Wt}};
    for (auto &o : out) {
        FILE *b = std::fopen((d + "/" + o.file).c_str(), "wb");
        if (!b) return 1;
        size_t cnt = (size_t)num_snps * 2 * o.W;
        size_t wr = std::fwrite(o.p, sizeof(uint32_t), cnt, b);
        std::fclose(b);
        if (wr != cnt) return 1;
    }
    return 0;
}

int episdet_read_header(const char *txt_path, int *num_snps, int *num_cases, int *num_controls, char *cases_path,
                        char *controls_path, int path_len)
{
    FILE *f = std::fopen(txt_path, "r");
    if (!f) return 1;
    char line[1000], fctl[1000], fcas[1000];
    bool ok = std::fgets(line, sizeof line, f) != nullptr;
    *num_snps = std::atoi(line);
    ok = ok && std::fgets(fctl, sizeof fctl, f) != nullptr;
    fctl[std::strcspn(fctl, "\r\n")] = 0;
    ok = ok && std::fgets(line, sizeof line, f) != nullptr;
    *num_controls = std::atoi(line);
    ok = ok && std::fgets(fcas, sizeof fcas, f) != nullptr;
    fcas[std::strcspn(fcas, "\r\n")] = 0;
    ok = ok && std::fgets(line, sizeof line, f) != nullptr;
    *num_cases = std::atoi(line);
    std::fclose(f);
    if (!ok || *num_snps <= 0 || *num_cases <= 0 || *num_controls <= 0) return 1;
    char *dup = strdup(txt_path);
    std::string base = dirname(dup);
    std::free(dup);
    std::string pc = base + "/" + fcas, pt = base + "/" + fctl;
    if ((int)pc.size() >= path_len || (int)pt.size() >= path_len) return 1;
    std::strcpy(cases_path, pc.c_str());
    std::strcpy(controls_path, pt.c_str());
    return 0;
}

int episdet_read_planes(const char *path, int num_snps, int num_samples, int num_snps_alloc, uint32_t *planes)
{
    const int W = episdet_padded_words(num_samples);
    FILE *f = std::fopen(path, "rb");
    if (!f) return 1;
    size_t want = (size_t)num_snps * 2 * W;
    size_t got = std::fread(planes, sizeof(uint32_t), want, f);
    std::fclose(f);
    if (num_snps_alloc > num_snps)
        std::memset(planes + want, 0, sizeof(uint32_t) * (size_t)(num_snps_alloc - num_snps) * 2 * W);
    return got == want ? 0 : 1;
}

}  // extern "C"
