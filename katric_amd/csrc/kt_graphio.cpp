// Host-side graph input layer: seeded synthetic generators, METIS and
// graph-io binary (.first_out / .head, uint64 little endian) readers and
// writers, CSR checksums.  Every generator returns a canonical symmetric CSR:
// uint32 ids, uint64 offsets, no self-loops, no duplicates, ascending lists.
//
// Reference counterpart: the graph-io library (graphio::read_graph,
// gen_local_graph; called from apps/shmetric.cpp:119, apps/katric.cpp:233-264),
// which is not vendored in the reference tree; the binary layout is the one of
// examples/metis-sample.first_out / .head.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/katric_graphio.h"

namespace {

// ---------------------------------------------------------------------------
// counter-based PRNG: SplitMix64 stream keyed by (seed, stream)
// ---------------------------------------------------------------------------
constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

inline uint64_t fmix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return fmix(seed * GOLDEN + fmix(stream + 0x632BE59BD9B4E019ull));
}
// i-th output of the SplitMix64 sequence with state `key`
inline uint64_t draw(uint64_t key, uint64_t i) { return fmix(key + (i + 1) * GOLDEN); }
// uniform in [0, n) from the high/low 32-bit half (n <= 2^32)
inline uint32_t below32(uint32_t r, uint64_t n) { return (uint32_t)(((uint64_t)r * n) >> 32); }

int num_threads() {
    const char* e = getenv("KATRIC_THREADS");
    if (!e) e = getenv("OMP_NUM_THREADS");  // the host's CPU share, where a scheduler sets it
    int t = e ? atoi(e) : (int)std::thread::hardware_concurrency();
    if (t < 1) t = 1;
    if (t > 64) t = 64;
    return t;
}

template <class F>
void parallel_for(int nt, F&& fn) {
    if (nt <= 1) {
        fn(0, 1);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t) th.emplace_back([&, t] { fn(t, nt); });
    for (auto& x : th) x.join();
}

inline void chunk(uint64_t len, int t, int nt, uint64_t& b, uint64_t& e) {
    b = len * t / nt;
    e = len * (t + 1) / nt;
}

// Parallel LSD radix sort of u64 keys over the low `bits` bits (8-bit digits).
void radix_sort(std::vector<uint64_t>& a, int bits, int nt) {
    const uint64_t len = a.size();
    if (len < 2) return;
    std::vector<uint64_t> b(len);
    std::vector<uint64_t> hist((size_t)nt * 256);
    int passes = (bits + 7) / 8;
    uint64_t* src = a.data();
    uint64_t* dst = b.data();
    for (int p = 0; p < passes; ++p) {
        const int sh = 8 * p;
        std::fill(hist.begin(), hist.end(), 0);
        parallel_for(nt, [&](int t, int n_t) {
            uint64_t lo, hi;
            chunk(len, t, n_t, lo, hi);
            uint64_t* h = &hist[(size_t)t * 256];
            for (uint64_t i = lo; i < hi; ++i) h[(src[i] >> sh) & 255]++;
        });
        uint64_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < nt; ++t) {
                uint64_t c = hist[(size_t)t * 256 + d];
                hist[(size_t)t * 256 + d] = run;
                run += c;
            }
        parallel_for(nt, [&](int t, int n_t) {
            uint64_t lo, hi;
            chunk(len, t, n_t, lo, hi);
            uint64_t* h = &hist[(size_t)t * 256];
            for (uint64_t i = lo; i < hi; ++i) dst[h[(src[i] >> sh) & 255]++] = src[i];
        });
        std::swap(src, dst);
    }
    if (src != a.data()) memcpy(a.data(), src, len * sizeof(uint64_t));
}

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

// Directed keys (u<<32|v) of both orientations -> canonical symmetric CSR.
int keys_to_csr(std::vector<uint64_t>& keys, uint32_t n, int nt, kt_host_graph* out) {
    radix_sort(keys, 32 + bits_for(n ? n - 1 : 0), nt);
    uint64_t w = 0;
    for (uint64_t i = 0; i < keys.size(); ++i)
        if (i == 0 || keys[i] != keys[i - 1]) keys[w++] = keys[i];
    keys.resize(w);
    out->n = n;
    out->nnz = w;
    out->offsets = (uint64_t*)malloc(((size_t)n + 1) * sizeof(uint64_t));
    out->heads = (uint32_t*)malloc((w ? w : 1) * sizeof(uint32_t));
    if (!out->offsets || !out->heads) {
        free(out->offsets);
        free(out->heads);
        out->offsets = nullptr;
        out->heads = nullptr;
        return KT_GIO_ENOMEM;
    }
    std::fill(out->offsets, out->offsets + (size_t)n + 1, 0);
    for (uint64_t i = 0; i < w; ++i) {
        out->offsets[(keys[i] >> 32) + 1]++;
        out->heads[i] = (uint32_t)keys[i];
    }
    for (uint64_t v = 0; v < n; ++v) out->offsets[v + 1] += out->offsets[v];
    return KT_GIO_OK;
}

// Emits the directed pair keys of accepted samples; sample(i, &u, &v) returns
// false for rejected samples.  Deterministic for any thread count.
template <class S>
void sample_keys(uint64_t samples, int nt, S&& sample, std::vector<uint64_t>& keys) {
    std::vector<uint64_t> cnt((size_t)nt + 1, 0);
    parallel_for(nt, [&](int t, int n_t) {
        uint64_t lo, hi, c = 0;
        chunk(samples, t, n_t, lo, hi);
        for (uint64_t i = lo; i < hi; ++i) {
            uint32_t u, v;
            if (sample(i, u, v) && u != v) c += 2;
        }
        cnt[t + 1] = c;
    });
    for (int t = 0; t < nt; ++t) cnt[t + 1] += cnt[t];
    keys.resize(cnt[nt]);
    parallel_for(nt, [&](int t, int n_t) {
        uint64_t lo, hi, p = cnt[t];
        chunk(samples, t, n_t, lo, hi);
        for (uint64_t i = lo; i < hi; ++i) {
            uint32_t u, v;
            if (sample(i, u, v) && u != v) {
                keys[p++] = ((uint64_t)u << 32) | v;
                keys[p++] = ((uint64_t)v << 32) | u;
            }
        }
    });
}

// G(n, m) / R-MAT into a canonical CSR using nt threads.
int gen_gnm_impl(uint32_t n, uint64_t samples, uint64_t seed, int nt, kt_host_graph* out) {
    memset(out, 0, sizeof(*out));
    if (n == 0) return KT_GIO_EINVAL;
    const uint64_t key = stream_key(seed, 0);
    std::vector<uint64_t> keys;
    sample_keys(samples, nt,
                [&](uint64_t i, uint32_t& u, uint32_t& v) {
                    uint64_t r = draw(key, i);
                    u = below32((uint32_t)r, n);
                    v = below32((uint32_t)(r >> 32), n);
                    return true;
                },
                keys);
    return keys_to_csr(keys, n, nt, out);
}

int gen_rmat_impl(uint32_t scale, uint64_t samples, double a, double b, double c, uint64_t seed, uint32_t n_limit,
                  int permute, int nt, kt_host_graph* out) {
    memset(out, 0, sizeof(*out));
    if (scale == 0 || scale > 32 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return KT_GIO_EINVAL;
    const uint64_t full = 1ull << scale;
    const uint64_t n64 = n_limit ? std::min<uint64_t>(n_limit, full) : full;
    if (n64 > 0xffffffffull) return KT_GIO_EINVAL;
    const uint32_t n = (uint32_t)n64;
    // seeded Fisher-Yates permutation of the surviving ids
    std::vector<uint32_t> perm;
    if (permute) {
        perm.resize(n);
        for (uint32_t i = 0; i < n; ++i) perm[i] = i;
        const uint64_t pk = stream_key(seed, 2);
        for (uint64_t i = n; i > 1; --i) {
            uint64_t j = (uint64_t)(((unsigned __int128)draw(pk, i) * i) >> 64);
            std::swap(perm[i - 1], perm[j]);
        }
    }
    const uint64_t key = stream_key(seed, 1);
    const double ab = a + b, abc = a + b + c;
    std::vector<uint64_t> keys;
    sample_keys(samples, nt,
                [&](uint64_t i, uint32_t& u, uint32_t& v) {
                    uint64_t x = 0, y = 0;
                    for (uint32_t l = 0; l < scale; ++l) {
                        double r = (double)(draw(key, i * scale + l) >> 11) * 0x1.0p-53;
                        uint64_t bx = 0, by = 0;
                        if (r < a) {
                        } else if (r < ab) {
                            by = 1;
                        } else if (r < abc) {
                            bx = 1;
                        } else {
                            bx = 1;
                            by = 1;
                        }
                        x = (x << 1) | bx;
                        y = (y << 1) | by;
                    }
                    if (x >= n64 || y >= n64) return false;
                    u = permute ? perm[x] : (uint32_t)x;
                    v = permute ? perm[y] : (uint32_t)y;
                    return true;
                },
                keys);
    return keys_to_csr(keys, n, nt, out);
}

}  // namespace

extern "C" {

void kt_host_graph_free(kt_host_graph* g) {
    if (!g) return;
    free(g->offsets);
    free(g->heads);
    g->offsets = nullptr;
    g->heads = nullptr;
    g->n = 0;
    g->nnz = 0;
}

int kt_gen_gnm(uint32_t n, uint64_t samples, uint64_t seed, kt_host_graph* out) {
    if (!out) return KT_GIO_EINVAL;
    return gen_gnm_impl(n, samples, seed, num_threads(), out);
}

int kt_gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c, uint64_t seed, uint32_t n_limit,
                int permute, kt_host_graph* out) {
    if (!out) return KT_GIO_EINVAL;
    return gen_rmat_impl(scale, samples, a, b, c, seed, n_limit, permute, num_threads(), out);
}

void kt_host_batch_free(kt_host_batch* b) {
    if (!b) return;
    free(b->vertex_base);
    free(b->offsets);
    free(b->heads);
    free(b->kind);
    memset(b, 0, sizeof(*b));
}

// Mixed batch: graph g draws, from its own SplitMix64 stream, n =
// floor(exp(U[ln n_min, ln n_max])), an average degree d ~ U[deg_min, deg_max]
// and its kind (R-MAT with probability rmat_fraction, else G(n,m)); both kinds
// take n*d/2 endpoint pairs; R-MAT uses scale ceil(log2 n), rejects ids >= n
// and permutes ids.  Output: one disjoint-union CSR with global ids.
int kt_gen_batch(uint32_t num_graphs, uint64_t seed, uint32_t n_min, uint32_t n_max, double deg_min, double deg_max,
                 double rmat_fraction, double a, double b, double c, kt_host_batch* out) {
    if (!out) return KT_GIO_EINVAL;
    memset(out, 0, sizeof(*out));
    if (num_graphs == 0 || n_min < 2 || n_max < n_min || deg_min < 0 || deg_max < deg_min) return KT_GIO_EINVAL;
    std::vector<kt_host_graph> gs(num_graphs);
    std::vector<uint8_t> kind(num_graphs);
    std::atomic<uint32_t> next{0};
    std::atomic<int> fail{0};
    const int nt = num_threads();
    parallel_for(nt, [&](int, int) {
        while (true) {
            uint32_t g = next.fetch_add(1);
            if (g >= num_graphs) break;
            const uint64_t k = stream_key(seed, 1000 + g);
            const double u0 = (double)(draw(k, 0) >> 11) * 0x1.0p-53;
            const double u1 = (double)(draw(k, 1) >> 11) * 0x1.0p-53;
            const double u2 = (double)(draw(k, 2) >> 11) * 0x1.0p-53;
            const double ln0 = std::log((double)n_min), ln1 = std::log((double)n_max);
            uint32_t n = (uint32_t)std::floor(std::exp(ln0 + u0 * (ln1 - ln0)));
            n = std::max(n_min, std::min(n, n_max));
            const double d = deg_min + u1 * (deg_max - deg_min);
            const uint64_t samples = (uint64_t)std::llround(n * d / 2.0);
            const uint64_t gseed = draw(k, 3);
            int rc;
            if (u2 < rmat_fraction) {
                kind[g] = 1;
                uint32_t scale = 1;
                while ((1ull << scale) < n) ++scale;
                rc = gen_rmat_impl(scale, samples, a, b, c, gseed, n, 1, 1, &gs[g]);
            } else {
                kind[g] = 0;
                rc = gen_gnm_impl(n, samples, gseed, 1, &gs[g]);
            }
            if (rc) fail.store(rc);
        }
    });
    if (fail.load()) {
        for (auto& g : gs) kt_host_graph_free(&g);
        return fail.load();
    }
    uint64_t N = 0, E = 0;
    for (auto& g : gs) {
        N += g.n;
        E += g.nnz;
    }
    if (N >= 0xffffffffull) {
        for (auto& g : gs) kt_host_graph_free(&g);
        return KT_GIO_EINVAL;
    }
    out->num_graphs = num_graphs;
    out->n_total = N;
    out->nnz_total = E;
    out->vertex_base = (uint64_t*)malloc(((size_t)num_graphs + 1) * 8);
    out->offsets = (uint64_t*)malloc((N + 1) * 8);
    out->heads = (uint32_t*)malloc((E ? E : 1) * 4);
    out->kind = (uint8_t*)malloc(num_graphs);
    if (!out->vertex_base || !out->offsets || !out->heads || !out->kind) {
        for (auto& g : gs) kt_host_graph_free(&g);
        kt_host_batch_free(out);
        return KT_GIO_ENOMEM;
    }
    memcpy(out->kind, kind.data(), num_graphs);
    std::vector<uint64_t> ebase(num_graphs + 1, 0);
    out->vertex_base[0] = 0;
    for (uint32_t g = 0; g < num_graphs; ++g) {
        out->vertex_base[g + 1] = out->vertex_base[g] + gs[g].n;
        ebase[g + 1] = ebase[g] + gs[g].nnz;
    }
    next = 0;
    parallel_for(nt, [&](int, int) {
        while (true) {
            uint32_t g = next.fetch_add(1);
            if (g >= num_graphs) break;
            const uint64_t vb = out->vertex_base[g], eb = ebase[g];
            for (uint32_t v = 0; v < gs[g].n; ++v) out->offsets[vb + v] = eb + gs[g].offsets[v];
            for (uint64_t e = 0; e < gs[g].nnz; ++e) out->heads[eb + e] = (uint32_t)(vb + gs[g].heads[e]);
            kt_host_graph_free(&gs[g]);
        }
    });
    out->offsets[N] = E;
    return KT_GIO_OK;
}

// METIS text: header "n m [fmt [ncon]]", then one 1-based adjacency line per
// vertex; '%' lines are comments.  fmt digits: sizes / vertex weights / edge
// weights (weights are skipped).
static int read_metis_impl(const char* path, kt_host_graph* out, FILE*& f) {
    f = fopen(path, "rb");
    if (!f) return KT_GIO_EIO;
    // every head needs at least two bytes of the file ("1 "): the header's
    // edge count only bounds the reservation, it is not trusted
    uint64_t fsize = 0;
    if (fseek(f, 0, SEEK_END) == 0) {
        const long e = ftell(f);
        fsize = e > 0 ? (uint64_t)e : 0;
    }
    rewind(f);
    std::string line;
    auto next_line = [&](std::string& s) -> bool {
        s.clear();
        int ch;
        bool any = false;
        while ((ch = fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            s.push_back((char)ch);
        }
        return any;
    };
    do {
        if (!next_line(line)) return KT_GIO_EFORMAT;
    } while (!line.empty() && line[0] == '%');
    unsigned long long n = 0, m = 0;
    char fmt[16] = "0";
    unsigned ncon = 1;
    int got = sscanf(line.c_str(), "%llu %llu %15s %u", &n, &m, fmt, &ncon);
    if (got < 2 || n > 0xffffffffull) return KT_GIO_EFORMAT;
    size_t fl = strlen(fmt);
    bool ew = fl >= 1 && fmt[fl - 1] == '1';
    bool vw = fl >= 2 && fmt[fl - 2] == '1';
    bool vs = fl >= 3 && fmt[fl - 3] == '1';
    if (!vw) ncon = 0;
    std::vector<uint64_t> off(n + 1, 0);
    std::vector<uint32_t> heads;
    heads.reserve(std::min<uint64_t>(2 * (uint64_t)m, fsize / 2 + 1));
    uint64_t v = 0;
    int rc = KT_GIO_OK;
    while (v < n && next_line(line)) {
        if (!line.empty() && line[0] == '%') continue;
        const char* p = line.c_str();
        char* endp;
        unsigned skip = (vs ? 1 : 0) + ncon;
        for (unsigned k = 0; k < skip; ++k) {
            strtoull(p, &endp, 10);
            p = endp;
        }
        while (true) {
            while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
            if (!*p) break;
            errno = 0;
            unsigned long long x = strtoull(p, &endp, 10);
            if (endp == p || errno) {
                rc = KT_GIO_EFORMAT;
                break;
            }
            p = endp;
            if (x == 0 || x > n) {
                rc = KT_GIO_EFORMAT;
                break;
            }
            heads.push_back((uint32_t)(x - 1));
            if (ew) {
                strtoull(p, &endp, 10);
                p = endp;
            }
        }
        if (rc) break;
        off[v + 1] = heads.size();
        ++v;
    }
    fclose(f);
    f = nullptr;
    if (rc) return rc;
    for (; v < n; ++v) off[v + 1] = heads.size();
    out->n = (uint32_t)n;
    out->nnz = heads.size();
    out->offsets = (uint64_t*)malloc((n + 1) * sizeof(uint64_t));
    out->heads = (uint32_t*)malloc((heads.size() ? heads.size() : 1) * sizeof(uint32_t));
    if (!out->offsets || !out->heads) {
        kt_host_graph_free(out);
        return KT_GIO_ENOMEM;
    }
    memcpy(out->offsets, off.data(), (n + 1) * sizeof(uint64_t));
    memcpy(out->heads, heads.data(), heads.size() * sizeof(uint32_t));
    return KT_GIO_OK;
}

// No C++ exception leaves the C ABI: allocation failures (a header that
// promises more vertices than memory holds) become KT_GIO_ENOMEM.
int kt_read_metis(const char* path, kt_host_graph* out) {
    if (!path || !out) return KT_GIO_EINVAL;
    memset(out, 0, sizeof(*out));
    FILE* f = nullptr;
    int rc;
    try {
        rc = read_metis_impl(path, out, f);
    } catch (const std::bad_alloc&) {
        rc = KT_GIO_ENOMEM;
    } catch (...) {
        rc = KT_GIO_EFORMAT;
    }
    if (f) fclose(f);
    if (rc) kt_host_graph_free(out);
    return rc;
}

int kt_write_metis(const char* path, uint32_t n, const uint64_t* offsets, const uint32_t* heads) {
    if (!path || !offsets || (n && !heads)) return KT_GIO_EINVAL;
    FILE* f = fopen(path, "wb");
    if (!f) return KT_GIO_EIO;
    fprintf(f, "%u %llu\n", n, (unsigned long long)(offsets[n] / 2));
    for (uint32_t v = 0; v < n; ++v) {
        for (uint64_t e = offsets[v]; e < offsets[v + 1]; ++e)
            fprintf(f, e == offsets[v] ? "%u" : " %u", heads[e] + 1);
        fputc('\n', f);
    }
    return fclose(f) == 0 ? KT_GIO_OK : KT_GIO_EIO;
}

static int read_graphio_binary_impl(const char* first_out_path, const char* head_path, kt_host_graph* out) {
    auto slurp = [](const char* p, std::vector<uint64_t>& v) -> int {
        FILE* f = fopen(p, "rb");
        if (!f) return KT_GIO_EIO;
        fseek(f, 0, SEEK_END);
        long sz = ftell(f);
        fseek(f, 0, SEEK_SET);
        if (sz < 0 || sz % 8) {
            fclose(f);
            return KT_GIO_EFORMAT;
        }
        v.resize((size_t)sz / 8);
        size_t r = v.empty() ? 0 : fread(v.data(), 8, v.size(), f);
        fclose(f);
        return r == v.size() ? KT_GIO_OK : KT_GIO_EIO;
    };
    std::vector<uint64_t> fo, hd;
    int rc = slurp(first_out_path, fo);
    if (rc) return rc;
    rc = slurp(head_path, hd);
    if (rc) return rc;
    if (fo.empty() || fo[0] != 0 || fo.back() != hd.size() || fo.size() - 1 > 0xffffffffull) return KT_GIO_EFORMAT;
    const uint64_t n = fo.size() - 1;
    for (uint64_t v = 0; v < n; ++v)
        if (fo[v + 1] < fo[v]) return KT_GIO_EFORMAT;
    for (uint64_t h : hd)
        if (h >= n) return KT_GIO_EFORMAT;
    out->n = (uint32_t)n;
    out->nnz = hd.size();
    out->offsets = (uint64_t*)malloc(fo.size() * sizeof(uint64_t));
    out->heads = (uint32_t*)malloc((hd.size() ? hd.size() : 1) * sizeof(uint32_t));
    if (!out->offsets || !out->heads) {
        kt_host_graph_free(out);
        return KT_GIO_ENOMEM;
    }
    memcpy(out->offsets, fo.data(), fo.size() * sizeof(uint64_t));
    for (size_t i = 0; i < hd.size(); ++i) out->heads[i] = (uint32_t)hd[i];
    return KT_GIO_OK;
}

int kt_read_graphio_binary(const char* first_out_path, const char* head_path, kt_host_graph* out) {
    if (!first_out_path || !head_path || !out) return KT_GIO_EINVAL;
    memset(out, 0, sizeof(*out));
    int rc;
    try {
        rc = read_graphio_binary_impl(first_out_path, head_path, out);
    } catch (const std::bad_alloc&) {
        rc = KT_GIO_ENOMEM;
    } catch (...) {
        rc = KT_GIO_EFORMAT;
    }
    if (rc) kt_host_graph_free(out);
    return rc;
}

int kt_write_graphio_binary(const char* first_out_path, const char* head_path, uint32_t n, const uint64_t* offsets,
                            const uint32_t* heads) {
    if (!first_out_path || !head_path || !offsets || (n && !heads)) return KT_GIO_EINVAL;
    FILE* f = fopen(first_out_path, "wb");
    if (!f) return KT_GIO_EIO;
    size_t w = fwrite(offsets, 8, (size_t)n + 1, f);
    if (fclose(f) != 0 || w != (size_t)n + 1) return KT_GIO_EIO;
    f = fopen(head_path, "wb");
    if (!f) return KT_GIO_EIO;
    for (uint64_t i = 0; i < offsets[n]; ++i) {
        uint64_t h = heads[i];
        if (fwrite(&h, 8, 1, f) != 1) {
            fclose(f);
            return KT_GIO_EIO;
        }
    }
    return fclose(f) == 0 ? KT_GIO_OK : KT_GIO_EIO;
}

// FNV-1a 64 over n, the offsets and the heads (little-endian words).
uint64_t kt_csr_checksum(uint32_t n, const uint64_t* offsets, const uint32_t* heads) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto eat = [&](const void* p, size_t len) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < len; ++i) {
            h ^= b[i];
            h *= 0x100000001b3ull;
        }
    };
    eat(&n, 4);
    if (offsets) eat(offsets, ((size_t)n + 1) * 8);
    if (offsets && heads) eat(heads, offsets[n] * 4);
    return h;
}

}  // extern "C"
