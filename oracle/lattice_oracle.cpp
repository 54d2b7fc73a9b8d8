// CPU restatement of ImageStack's permutohedral-lattice Gauss transform.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke check
// and bench.py's cpu_baseline leg call it to check results (and as the scalar
// CPU baseline).  The product path (imagestack_amd/libgtlattice.so) never
// links or calls it.
//
// It restates, in our own code, the semantics of
//   src/Permutohedral.h            (lattice: hash table, splat, blur, slice)
//   src/GaussTransform.cpp:115-226 (GaussTransform::apply, PERMUTOHEDRAL case)
//   src/GaussTransform.cpp:504-678 (JointBilateral::apply, lattice branch)
//   src/Expr.h:2439-2458           (expr / float is expr * (1.0f/float))
// of the reference, and must be compiled strictly: -O2 -ffp-contract=off
// -fno-fast-math (no FMA contraction, IEEE single rounding per operation).
//
// Three arithmetic modes:
//   ORC_F32       : FP32 everywhere, the reference's own arithmetic order.  On
//                   the same inputs its output is bit-identical to the
//                   reference built with strict flags (pinned by tests that
//                   compare against oracle/_ref and committed fixtures).
//   ORC_F64_LAT   : "FP64 oracle of the same lattice": keys, ranks and
//                   barycentric weights come from the FP32 elevation exactly as
//                   the reference computes them; splat sums, blur, slice and
//                   normalisation run in FP64.  This is the 1e-5 gate.
//   ORC_F64_PURE  : keys and ranks from the FP32 elevation (so the key set is
//                   shared) but the elevation and barycentric weights are
//                   recomputed in FP64.  Reported, not gated.
//
// Parity pins for each function are listed in tests/test_oracle.py.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <numeric>
#include <vector>

#include "grid_oracle.h"

namespace {

enum { ORC_F32 = 0, ORC_F64_LAT = 1, ORC_F64_PURE = 2 };
constexpr int kMaxD = 64;

// --- lattice constants ----------------------------------------------------
// src/Permutohedral.h:257-274.  Note the reference feeds sqrtf a double
// (2.0/3), which C converts to float before the call.
void scale_factors_f32(int d, float *sf) {
    const float s23 = sqrtf((float)(2.0 / 3));
    for (int i = 0; i < d; i++) {
        float v = 1.0f / (sqrtf((float)(i + 1) * (float)(i + 2)));
        v *= (float)(d + 1) * s23;
        sf[i] = v;
    }
}

void scale_factors_f64(int d, double *sf) {
    for (int i = 0; i < d; i++)
        sf[i] = (1.0 / std::sqrt((double)(i + 1) * (double)(i + 2))) *
                ((double)(d + 1) * std::sqrt(2.0 / 3.0));
}

// --- enclosing simplex ----------------------------------------------------
// src/Permutohedral.h:293-356.  Fills greedy (fixed-up), rank, the FP32
// elevation and the FP32 barycentric weights b[0..d].
struct Simplex {
    int greedy[kMaxD + 1];
    int rank[kMaxD + 1];
    float elev[kMaxD + 1];
    float bary[kMaxD + 2];
    bool key_overflow;
};

void simplex_f32(int d, const float *p, const float *sf, Simplex &s) {
    float *e = s.elev;
    // Elevation (:293-298): products associate left to right, the integer
    // factor is converted to float first, no contraction.
    e[d] = ((float)(-d) * p[d - 1]) * sf[d - 1];
    for (int i = d - 1; i > 0; i--)
        e[i] = (e[i + 1] - ((float)i * p[i - 1]) * sf[i - 1]) +
               ((float)(i + 2) * p[i]) * sf[i];
    e[0] = e[1] + (2.0f * p[0]) * sf[0];

    // Greedy rounding to the nearest remainder-0 point (:300-317).  Ties go
    // down because the comparison is strict.
    const float scale = 1.0f / (float)(d + 1);
    int sum = 0;
    s.key_overflow = false;
    for (int i = 0; i <= d; i++) {
        float v = e[i] * scale;
        float up = ceilf(v) * (float)(d + 1);
        float down = floorf(v) * (float)(d + 1);
        float g = (up - e[i] < e[i] - down) ? up : down;
        if (!(fabsf(g) <= 32767.0f - (float)(2 * (d + 1)))) s.key_overflow = true;
        s.greedy[i] = (int)(short)(int)g;
        sum += s.greedy[i];
    }
    sum /= (d + 1);

    // Rank differential (:321-324): ties increment the later index.
    for (int i = 0; i <= d; i++) s.rank[i] = 0;
    for (int i = 0; i < d; i++)
        for (int j = i + 1; j <= d; j++)
            if (e[i] - (float)s.greedy[i] < e[j] - (float)s.greedy[j]) s.rank[i]++;
            else s.rank[j]++;

    // Return to the hyperplane (:326-348).
    if (sum > 0) {
        for (int i = 0; i <= d; i++) {
            if (s.rank[i] >= d + 1 - sum) { s.greedy[i] -= d + 1; s.rank[i] += sum - (d + 1); }
            else s.rank[i] += sum;
        }
    } else if (sum < 0) {
        for (int i = 0; i <= d; i++) {
            if (s.rank[i] < -sum) { s.greedy[i] += d + 1; s.rank[i] += (d + 1) + sum; }
            else s.rank[i] += sum;
        }
    }

    // Barycentric weights (:350-356), accumulated in index order.
    for (int i = 0; i < d + 2; i++) s.bary[i] = 0.0f;
    for (int i = 0; i <= d; i++) {
        s.bary[d - s.rank[i]] += (e[i] - (float)s.greedy[i]) * scale;
        s.bary[d + 1 - s.rank[i]] -= (e[i] - (float)s.greedy[i]) * scale;
    }
    s.bary[0] += 1.0f + s.bary[d + 1];
}

// FP64 barycentric weights for the simplex chosen by the FP32 pass.
void bary_f64(int d, const float *p, const double *sf, const Simplex &s, double *b) {
    double e[kMaxD + 1];
    e[d] = (-(double)d * p[d - 1]) * sf[d - 1];
    for (int i = d - 1; i > 0; i--)
        e[i] = (e[i + 1] - ((double)i * p[i - 1]) * sf[i - 1]) + ((double)(i + 2) * p[i]) * sf[i];
    e[0] = e[1] + (2.0 * p[0]) * sf[0];
    const double scale = 1.0 / (double)(d + 1);
    for (int i = 0; i < d + 2; i++) b[i] = 0.0;
    for (int i = 0; i <= d; i++) {
        b[d - s.rank[i]] += (e[i] - (double)s.greedy[i]) * scale;
        b[d + 1 - s.rank[i]] -= (e[i] - (double)s.greedy[i]) * scale;
    }
    b[0] += 1.0 + b[d + 1];
}

// Vertex r of the simplex, first d coordinates (:365-369).  The canonical
// simplex entry is r if rank <= d-r, else r-(d+1) (:247-254).
inline void vertex_key(int d, const Simplex &s, int r, short *key) {
    for (int i = 0; i < d; i++) {
        int c = (s.rank[i] <= d - r) ? r : r - (d + 1);
        key[i] = (short)(s.greedy[i] + c);
    }
}

// --- insertion-ordered open-addressing table ------------------------------
// Same contract as HashTablePermutohedral (src/Permutohedral.h:18-164): keys
// of d shorts, linear probing, doubling when half full, keys and values in
// insertion order, values zero-initialised.  The hash function is our own
// (the reference's is not part of the contract: only the key set is).
template <typename T>
struct Table {
    int kd, vd;
    size_t capacity = 0, filled = 0;
    std::vector<int> entries;       // slot -> vertex index, -1 empty
    std::vector<short> keys;        // filled * kd
    std::vector<T> values;          // filled * vd

    Table(int kd_, int vd_) : kd(kd_), vd(vd_) {
        capacity = 1 << 15;
        entries.assign(capacity, -1);
        keys.reserve((capacity / 2) * kd);
        values.reserve((capacity / 2) * vd);
    }

    static uint64_t hash(const short *k, int kd) {
        uint64_t h = 0xcbf29ce484222325ull;
        for (int i = 0; i < kd; i++) {
            h ^= (uint16_t)k[i];
            h *= 0x100000001b3ull;
        }
        return h ^ (h >> 29);
    }

    bool equal(int idx, const short *k) const {
        const short *kk = &keys[(size_t)idx * kd];
        for (int i = 0; i < kd; i++) if (kk[i] != k[i]) return false;
        return true;
    }

    void grow() {
        size_t nc = capacity * 2;
        std::vector<int> ne(nc, -1);
        for (size_t i = 0; i < capacity; i++) {
            int idx = entries[i];
            if (idx < 0) continue;
            size_t h = hash(&keys[(size_t)idx * kd], kd) & (nc - 1);
            while (ne[h] != -1) h = (h + 1) & (nc - 1);
            ne[h] = idx;
        }
        entries.swap(ne);
        capacity = nc;
    }

    // Returns the vertex index, creating it when `create` (else -1).
    int lookup(const short *k, bool create) {
        if (create && filled >= capacity / 2 - 1) grow();
        size_t h = hash(k, kd) & (capacity - 1);
        while (true) {
            int idx = entries[h];
            if (idx < 0) {
                if (!create) return -1;
                idx = (int)filled++;
                entries[h] = idx;
                keys.insert(keys.end(), k, k + kd);
                values.insert(values.end(), (size_t)vd, T(0));
                return idx;
            }
            if (equal(idx, k)) return idx;
            h = (h + 1) & (capacity - 1);
        }
    }
};

// --- the lattice ------------------------------------------------------------
template <typename T>
struct Lattice {
    int d, vd, mode;
    float sf[kMaxD];
    double sf64[kMaxD];
    Table<T> table;
    std::vector<int> replay_idx;    // N*(d+1) vertex indices
    std::vector<T> replay_w;        // N*(d+1) weights
    bool overflow = false;

    Lattice(int d_, int vd_, int mode_) : d(d_), vd(vd_), mode(mode_), table(d_, vd_) {
        scale_factors_f32(d, sf);
        scale_factors_f64(d, sf64);
    }

    // Fills per-vertex weights w[0..d] in T precision, and the simplex.
    void weights(const float *p, Simplex &s, T *w) {
        simplex_f32(d, p, sf, s);
        if (s.key_overflow) overflow = true;
        if (mode == ORC_F64_PURE) {
            double b[kMaxD + 2];
            bary_f64(d, p, sf64, s, b);
            for (int r = 0; r <= d; r++) w[r] = (T)b[r];
        } else {
            for (int r = 0; r <= d; r++) w[r] = (T)s.bary[r];
        }
    }

    // src/Permutohedral.h:290-393 with splatting=true, record=true.
    void splat(const float *p, const T *val) {
        Simplex s;
        T w[kMaxD + 1];
        short key[kMaxD];
        weights(p, s, w);
        for (int r = 0; r <= d; r++) {
            vertex_key(d, s, r, key);
            int idx = table.lookup(key, true);
            T *v = &table.values[(size_t)idx * vd];
            for (int c = 0; c < vd; c++) v[c] += w[r] * val[c];
            replay_idx.push_back(idx);
            replay_w.push_back(w[r]);
        }
    }

    // src/Permutohedral.h:430-488: d+1 Jacobi passes of [1/4 1/2 1/4],
    // neighbours looked up per pass, missing neighbours are zero.
    void blur() {
        const size_t M = table.filled;
        std::vector<T> nv(M * vd), zero(vd, T(0));
        std::vector<short> n1(d), n2(d);
        std::vector<T> *oldv = &table.values, *newv = &nv;
        for (int j = 0; j <= d; j++) {
            for (size_t i = 0; i < M; i++) {
                const short *k = &table.keys[i * d];
                for (int c = 0; c < d; c++) { n1[c] = (short)(k[c] + 1); n2[c] = (short)(k[c] - 1); }
                if (j < d) { n1[j] = (short)(k[j] - d); n2[j] = (short)(k[j] + d); }
                int a = table.lookup(n1.data(), false);
                int b = table.lookup(n2.data(), false);
                const T *va = a >= 0 ? &(*oldv)[(size_t)a * vd] : zero.data();
                const T *vb = b >= 0 ? &(*oldv)[(size_t)b * vd] : zero.data();
                const T *vo = &(*oldv)[i * vd];
                T *vn = &(*newv)[i * vd];
                for (int c = 0; c < vd; c++)
                    vn[c] = (T(0.25) * va[c] + T(0.5) * vo[c]) + T(0.25) * vb[c];
            }
            std::swap(oldv, newv);
        }
        if (oldv != &table.values) table.values.swap(*oldv);
    }

    // Replay slice (src/Permutohedral.h:418-427).
    void slice_replay(size_t pix, T *col) const {
        for (int c = 0; c < vd; c++) col[c] = T(0);
        for (int r = 0; r <= d; r++) {
            int idx = replay_idx[pix * (d + 1) + r];
            T w = replay_w[pix * (d + 1) + r];
            const T *v = &table.values[(size_t)idx * vd];
            for (int c = 0; c < vd; c++) col[c] += w * v[c];
        }
    }

    // Non-replay slice (src/Permutohedral.h:405 via :290-393, splatting=false).
    // The reference inserts zero vertices for unseen keys here; they change
    // size() but not the output, so this lookup does not create them and
    // only counts them.
    size_t slice_created = 0;
    std::vector<std::vector<short>> seen_missing;
    void slice_pos(const float *p, T *col) {
        Simplex s;
        T w[kMaxD + 1];
        short key[kMaxD];
        weights(p, s, w);
        for (int c = 0; c < vd; c++) col[c] = T(0);
        for (int r = 0; r <= d; r++) {
            vertex_key(d, s, r, key);
            int idx = table.lookup(key, false);
            if (idx < 0) continue;
            const T *v = &table.values[(size_t)idx * vd];
            for (int c = 0; c < vd; c++) col[c] += w[r] * v[c];
        }
    }
};

struct PlanarIn {
    const float *base;
    int64_t cstride;
};

// GaussTransform::apply, PERMUTOHEDRAL case (src/GaussTransform.cpp:163-226)
// on dense planar images of n pixels in raster (t, y, x) order.
template <typename T>
void run_transform(Lattice<T> &lat, int n, const float *splat, int64_t splat_cs,
                   const float *values, int64_t values_cs, const float *inv_sigma,
                   bool do_blur) {
    const int d = lat.d, vd = lat.vd;
    lat.replay_idx.reserve((size_t)n * (d + 1));
    lat.replay_w.reserve((size_t)n * (d + 1));
    float pos[kMaxD];
    T val[kMaxD];
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < d; c++) pos[c] = splat[c * splat_cs + i] * inv_sigma[c];
        for (int c = 0; c < vd; c++) val[c] = (T)values[c * values_cs + i];
        lat.splat(pos, val);
    }
    if (do_blur) lat.blur();
}

template <typename T>
void write_out(Lattice<T> &lat, int n, int n_slice, const float *slice, int64_t slice_cs,
               const float *inv_sigma, T *out, int64_t out_cs) {
    const int d = lat.d, vd = lat.vd;
    T col[kMaxD];
    float pos[kMaxD];
    if (!slice) {
        for (int i = 0; i < n; i++) {
            lat.slice_replay(i, col);
            for (int c = 0; c < vd; c++) out[c * out_cs + i] = col[c];
        }
    } else {
        for (int i = 0; i < n_slice; i++) {
            for (int c = 0; c < d; c++) pos[c] = slice[c * slice_cs + i] * inv_sigma[c];
            lat.slice_pos(pos, col);
            for (int c = 0; c < vd; c++) out[c * out_cs + i] = col[c];
        }
    }
}

bool bad_dims(int d, int vd) { return d < 1 || d > kMaxD || vd < 1 || vd > kMaxD; }

// Sorted export of a lattice: lexicographic over signed int16 coordinates.
template <typename T>
std::vector<int> sorted_order(const Table<T> &t) {
    std::vector<int> ord(t.filled);
    std::iota(ord.begin(), ord.end(), 0);
    const int d = t.kd;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) {
        const short *ka = &t.keys[(size_t)a * d], *kb = &t.keys[(size_t)b * d];
        for (int i = 0; i < d; i++) if (ka[i] != kb[i]) return ka[i] < kb[i];
        return false;
    });
    return ord;
}

struct LatticeHandle {
    int mode, d, vd;
    int M;
    std::vector<short> keys;     // sorted, M*d
    std::vector<int> nbr;        // (d+1)*M*2, axis-major, sorted ids or -1
    std::vector<double> splat_v; // sorted, M*vd (values cast to double)
    std::vector<double> blur_v;  // sorted, M*vd
};

} // namespace

extern "C" {

void orc_scale_factors(int d, float *sf) { scale_factors_f32(d, sf); }

// Per-point enclosing simplex of one (already scaled) position.
// keys: (d+1) rows of d int16 (vertex r); w: d+1 FP32 weights; also greedy
// and rank (d+1 ints each) when non-NULL.  Returns 1 if a key would leave the
// int16 range, else 0.
int orc_simplex(int d, const float *pos, short *keys, float *w, int *greedy, int *rank) {
    float sf[kMaxD];
    scale_factors_f32(d, sf);
    Simplex s;
    simplex_f32(d, pos, sf, s);
    for (int r = 0; r <= d; r++) {
        vertex_key(d, s, r, keys + (size_t)r * d);
        w[r] = s.bary[r];
    }
    if (greedy) for (int i = 0; i <= d; i++) greedy[i] = s.greedy[i];
    if (rank) for (int i = 0; i <= d; i++) rank[i] = s.rank[i];
    return s.key_overflow ? 1 : 0;
}

// Batch form over planar positions: pos(c, i) = splat[c*cs + i] * inv_sigma[c]
// (src/GaussTransform.cpp:176-178).  keys: n*(d+1)*d, w: n*(d+1).
int orc_simplex_batch(int n, int d, const float *splat, int64_t cs, const float *inv_sigma,
                      short *keys, float *w) {
    if (bad_dims(d, 1)) return -1;
    float sf[kMaxD], pos[kMaxD];
    scale_factors_f32(d, sf);
    int ovf = 0;
    Simplex s;
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < d; c++) pos[c] = splat[c * cs + i] * inv_sigma[c];
        simplex_f32(d, pos, sf, s);
        ovf |= s.key_overflow;
        for (int r = 0; r <= d; r++) {
            vertex_key(d, s, r, keys + ((size_t)i * (d + 1) + r) * d);
            w[(size_t)i * (d + 1) + r] = s.bary[r];
        }
    }
    return ovf;
}

// GaussTransform::apply (PERMUTOHEDRAL) on dense planar images.
//   mode 0: out is float*, modes 1/2: out is double*.
//   slice == NULL selects the replay slice (slice image == splat image).
//   M_out receives the vertex count after splatting.
// Returns 0 on success, -1 on bad arguments, 1 if a key overflowed int16.
int orc_gauss_transform(int mode, int n, int d, int vd,
                        const float *splat, int64_t splat_cs,
                        const float *values, int64_t values_cs,
                        int n_slice, const float *slice, int64_t slice_cs,
                        const float *inv_sigma, void *out, int64_t out_cs, int *M_out) {
    if (bad_dims(d, vd) || n < 0 || mode < 0 || mode > 2) return -1;
    if (mode == ORC_F32) {
        Lattice<float> lat(d, vd, mode);
        run_transform(lat, n, splat, splat_cs, values, values_cs, inv_sigma, true);
        if (M_out) *M_out = (int)lat.table.filled;
        write_out(lat, n, n_slice, slice, slice_cs, inv_sigma, (float *)out, out_cs);
        return lat.overflow ? 1 : 0;
    } else {
        Lattice<double> lat(d, vd, mode);
        run_transform(lat, n, splat, splat_cs, values, values_cs, inv_sigma, true);
        if (M_out) *M_out = (int)lat.table.filled;
        write_out(lat, n, n_slice, slice, slice_cs, inv_sigma, (double *)out, out_cs);
        return lat.overflow ? 1 : 0;
    }
}

// Builds the lattice and exports it in sorted order.  The neighbour table is
// computed by binary search in the sorted key list from the closed form
// n1 = k + 1 - (d+1) e_j, n2 = k - 1 + (d+1) e_j (src/Permutohedral.h:446-451),
// independently of the hash table.
void *orc_lattice_build(int mode, int n, int d, int vd,
                        const float *splat, int64_t splat_cs,
                        const float *values, int64_t values_cs, const float *inv_sigma) {
    if (bad_dims(d, vd) || n < 0 || mode < 0 || mode > 2) return nullptr;
    LatticeHandle *h = new LatticeHandle();
    h->mode = mode; h->d = d; h->vd = vd;
    auto fill = [&](auto &lat) {
        run_transform(lat, n, splat, splat_cs, values, values_cs, inv_sigma, false);
        const int M = (int)lat.table.filled;
        h->M = M;
        std::vector<int> ord = sorted_order(lat.table);
        h->keys.resize((size_t)M * d);
        h->splat_v.resize((size_t)M * vd);
        for (int i = 0; i < M; i++) {
            memcpy(&h->keys[(size_t)i * d], &lat.table.keys[(size_t)ord[i] * d], sizeof(short) * d);
            for (int c = 0; c < vd; c++) h->splat_v[(size_t)i * vd + c] = (double)lat.table.values[(size_t)ord[i] * vd + c];
        }
        lat.blur();
        h->blur_v.resize((size_t)M * vd);
        for (int i = 0; i < M; i++)
            for (int c = 0; c < vd; c++) h->blur_v[(size_t)i * vd + c] = (double)lat.table.values[(size_t)ord[i] * vd + c];
    };
    if (mode == ORC_F32) { Lattice<float> lat(d, vd, mode); fill(lat); }
    else { Lattice<double> lat(d, vd, mode); fill(lat); }

    const int M = h->M;
    h->nbr.assign((size_t)(d + 1) * M * 2, -1);
    auto find = [&](const short *k) -> int {
        int lo = 0, hi = M;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            const short *km = &h->keys[(size_t)mid * d];
            int cmp = 0;
            for (int i = 0; i < d && !cmp; i++) cmp = (km[i] < k[i]) ? -1 : (km[i] > k[i]) ? 1 : 0;
            if (cmp == 0) return mid;
            if (cmp < 0) lo = mid + 1; else hi = mid;
        }
        return -1;
    };
    std::vector<short> n1(d), n2(d);
    for (int j = 0; j <= d; j++)
        for (int i = 0; i < M; i++) {
            const short *k = &h->keys[(size_t)i * d];
            for (int c = 0; c < d; c++) { n1[c] = (short)(k[c] + 1); n2[c] = (short)(k[c] - 1); }
            if (j < d) { n1[j] = (short)(k[j] - d); n2[j] = (short)(k[j] + d); }
            h->nbr[((size_t)j * M + i) * 2 + 0] = find(n1.data());
            h->nbr[((size_t)j * M + i) * 2 + 1] = find(n2.data());
        }
    return h;
}

int orc_lattice_size(void *handle) { return handle ? ((LatticeHandle *)handle)->M : -1; }

// keys: M*d int16; nbr: (d+1)*M*2 int32; splat_v, blur_v: M*vd double.
// Any output may be NULL.
void orc_lattice_export(void *handle, short *keys, int *nbr, double *splat_v, double *blur_v) {
    LatticeHandle *h = (LatticeHandle *)handle;
    if (keys) memcpy(keys, h->keys.data(), h->keys.size() * sizeof(short));
    if (nbr) memcpy(nbr, h->nbr.data(), h->nbr.size() * sizeof(int));
    if (splat_v) memcpy(splat_v, h->splat_v.data(), h->splat_v.size() * sizeof(double));
    if (blur_v) memcpy(blur_v, h->blur_v.data(), h->blur_v.size() * sizeof(double));
}

void orc_lattice_free(void *handle) { delete (LatticeHandle *)handle; }

// JointBilateral::apply, lattice branch (src/GaussTransform.cpp:504-678), on
// a dense planar image im (C channels) and guide ref (Cr channels) of
// W x H x F pixels.  Result written to out (C planes; float* for mode 0,
// double* otherwise).  Zero spatial sigmas filter columns, rows or frames
// independently (:516-538).
} // extern "C"

namespace {
struct View { const float *base; int W, H, F, C; int64_t ys, ts, cs; };

// JointBilateral::apply (src/GaussTransform.cpp:504-678) with the lattice
// (grid = false) or the bilateral grid (grid = true) as the Gauss transform.
template <typename T>
int jb_rec(int mode, View im, View ref, T *out, int64_t oys, int64_t ots, int64_t ocs,
                  float fw, float fh, float ff, float colorSigma, bool grid = false) {
    if (im.W == 1) fw = 0;
    if (im.H == 1) fh = 0;
    if (im.F == 1) ff = 0;
    if (im.W > 1 && fw == 0) {
        for (int x = 0; x < im.W; x++) {
            View a = im, b = ref; a.base += x; b.base += x; a.W = b.W = 1;
            int rc = jb_rec<T>(mode, a, b, out + x, oys, ots, ocs, 0, fh, ff, colorSigma, grid);
            if (rc < 0) return rc;
        }
        return 0;
    }
    if (im.H > 1 && fh == 0) {
        for (int y = 0; y < im.H; y++) {
            View a = im, b = ref; a.base += y * im.ys; b.base += y * ref.ys; a.H = b.H = 1;
            int rc = jb_rec<T>(mode, a, b, out + y * oys, oys, ots, ocs, fw, 0, ff, colorSigma, grid);
            if (rc < 0) return rc;
        }
        return 0;
    }
    if (im.F > 1 && ff == 0) {
        for (int t = 0; t < im.F; t++) {
            View a = im, b = ref; a.base += t * im.ts; b.base += t * ref.ts; a.F = b.F = 1;
            int rc = jb_rec<T>(mode, a, b, out + t * ots, oys, ots, ocs, fw, fh, 0, colorSigma, grid);
            if (rc < 0) return rc;
        }
        return 0;
    }
    const bool fx = im.W > 1 && fw < 10 * im.W;
    const bool fy = im.H > 1 && fh < 10 * im.H;
    const bool ft = im.F > 1 && ff < 10 * im.F;
    const int d = ref.C + (fx ? 1 : 0) + (fy ? 1 : 0) + (ft ? 1 : 0);
    const int vd = im.C + 1;
    if (bad_dims(d, vd)) return -1;
    const int n = im.W * im.H * im.F;
    // Positions: ref * (1.0f/colorSigma), then X, Y, T times (1.0f/sigma)
    // (:654-666 with the reciprocal rule of src/Expr.h:2439-2458).
    std::vector<float> pos((size_t)d * n), val((size_t)vd * n);
    const float ic = 1.0f / colorSigma, ix = 1.0f / fw, iy = 1.0f / fh, it = 1.0f / ff;
    size_t i = 0;
    for (int t = 0; t < im.F; t++)
        for (int y = 0; y < im.H; y++)
            for (int x = 0; x < im.W; x++, i++) {
                int c = 0;
                for (; c < ref.C; c++) pos[(size_t)c * n + i] = ref.base[c * ref.cs + t * ref.ts + y * ref.ys + x] * ic;
                if (fx) pos[(size_t)(c++) * n + i] = (float)x * ix;
                if (fy) pos[(size_t)(c++) * n + i] = (float)y * iy;
                if (ft) pos[(size_t)(c++) * n + i] = (float)t * it;
                for (int k = 0; k < im.C; k++) val[(size_t)k * n + i] = im.base[k * im.cs + t * im.ts + y * im.ys + x];
                val[(size_t)im.C * n + i] = 1.0f;
            }
    std::vector<float> ones(d, 1.0f / 1.0f);
    std::vector<T> res((size_t)vd * n);
    bool overflow = false;
    if (grid) {
        grid_oracle::transform<T>(n, d, vd, pos.data(), n, val.data(), n, 0, nullptr, 0, ones.data(), res.data(), n);
    } else {
        Lattice<T> lat(d, vd, mode);
        run_transform(lat, n, pos.data(), n, val.data(), n, ones.data(), true);
        write_out(lat, n, n, nullptr, 0, ones.data(), res.data(), n);
        overflow = lat.overflow;
    }
    // Normalise by the homogeneous channel, true division (:673-676).
    i = 0;
    for (int t = 0; t < im.F; t++)
        for (int y = 0; y < im.H; y++)
            for (int x = 0; x < im.W; x++, i++)
                for (int k = 0; k < im.C; k++)
                    out[k * ocs + t * ots + y * oys + x] = res[(size_t)k * n + i] / res[(size_t)im.C * n + i];
    return overflow ? 1 : 0;
}
} // namespace

extern "C" {
int orc_joint_bilateral(int mode, int W, int H, int F, int C, const float *im,
                        int Cr, const float *ref, float fw, float fh, float ff,
                        float colorSigma, void *out) {
    if (W < 1 || H < 1 || F < 1 || C < 1 || Cr < 1 || mode < 0 || mode > 2) return -1;
    const int64_t ys = W, ts = (int64_t)W * H, cs = (int64_t)W * H * F;
    View a{im, W, H, F, C, ys, ts, cs}, b{ref, W, H, F, Cr, ys, ts, cs};
    if (mode == ORC_F32) return jb_rec<float>(mode, a, b, (float *)out, ys, ts, cs, fw, fh, ff, colorSigma);
    return jb_rec<double>(mode, a, b, (double *)out, ys, ts, cs, fw, fh, ff, colorSigma);
}

// JointBilateral::apply(..., GRID): the same positions and values, the
// bilateral grid as the Gauss transform (mode F32 or F64_LAT: FP64 weights,
// sums, blur and slice on the FP32 cell indices).
int orc_joint_bilateral_grid(int mode, int W, int H, int F, int C, const float *im,
                             int Cr, const float *ref, float fw, float fh, float ff,
                             float colorSigma, void *out) {
    if (W < 1 || H < 1 || F < 1 || C < 1 || Cr < 1 || mode < 0 || mode > 2) return -1;
    const int64_t ys = W, ts = (int64_t)W * H, cs = (int64_t)W * H * F;
    View a{im, W, H, F, C, ys, ts, cs}, b{ref, W, H, F, Cr, ys, ts, cs};
    if (mode == ORC_F32) return jb_rec<float>(mode, a, b, (float *)out, ys, ts, cs, fw, fh, ff, colorSigma, true);
    return jb_rec<double>(mode, a, b, (double *)out, ys, ts, cs, fw, fh, ff, colorSigma, true);
}

// GaussTransform::apply(slice, splat, values, sigma, GRID)
// (src/GaussTransform.cpp:227-297); slice == NULL slices at the splat
// positions.  Returns the grid's cell count (or -1 on bad dimensions).
int64_t orc_grid_transform(int mode, int n, int d, int vd, const float *splat, int64_t splat_cs,
                           const float *values, int64_t values_cs, int n_slice, const float *slice,
                           int64_t slice_cs, const float *inv_sigma, void *out, int64_t out_cs) {
    if (d < 1 || d > 8 || vd < 1 || vd > 16 || n < 0 || mode < 0 || mode > 2) return -1;
    if (mode == ORC_F32)
        return grid_oracle::transform<float>(n, d, vd, splat, splat_cs, values, values_cs, n_slice, slice,
                                             slice_cs, inv_sigma, (float *)out, out_cs);
    return grid_oracle::transform<double>(n, d, vd, splat, splat_cs, values, values_cs, n_slice, slice,
                                          slice_cs, inv_sigma, (double *)out, out_cs);
}

} // extern "C"

// --- non-local means (NLMeans::apply, src/GaussTransform.cpp:867-874 and
// NLMeans3D::apply, :934-941) after the PCA ---------------------------------
namespace {

// Convolve::apply(im, filters, Zero, Multiply::Inner) (src/Convolve.cpp:248-270)
// for a bank of K*C filter channels (PatchPCA's output, src/Statistics.cpp:
// 1710-1770): output channel k is the sum over c = 0..C-1, in that order, of
// convolveSingle(image channel c, filter channel k*C + c) with the zero
// boundary (:151-173): a true convolution (the filter is read at xoff-dx,
// yoff-dy, toff-dt), taps outside the image skipped, taps summed in t, y, x
// order into v = 0, then added to the output.  T = float is the reference's
// own arithmetic; T = double sums in FP64.
template <typename T>
void patch_project(int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int K,
                   const float *filt, T *out) {
    const int64_t plane = (int64_t)W * H, n = plane * F;
    const int64_t fsz = (int64_t)fw * fh * ft;
    const int xoff = (fw - 1) / 2, yoff = (fh - 1) / 2, toff = (ft - 1) / 2;
    for (int64_t i = 0; i < (int64_t)K * n; i++) out[i] = 0;
    for (int i = 0; i < K * C; i++) {
        const float *src = im + (int64_t)(i % C) * n;
        const float *w = filt + (int64_t)i * fsz;
        T *dst = out + (int64_t)(i / C) * n;
        for (int t = 0; t < F; t++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    T v = 0;
                    for (int dt = -toff; dt <= toff; dt++) {
                        if (t + dt < 0) continue;
                        if (t + dt >= F) break;
                        for (int dy = -yoff; dy <= yoff; dy++) {
                            if (y + dy < 0) continue;
                            if (y + dy >= H) break;
                            for (int dx = -xoff; dx <= xoff; dx++) {
                                if (x + dx < 0) continue;
                                if (x + dx >= W) break;
                                const T wv = (T)w[(xoff - dx) + fw * ((yoff - dy) + (int64_t)fh * (toff - dt))];
                                v += (T)src[(t + dt) * plane + (int64_t)(y + dy) * W + (x + dx)] * wv;
                            }
                        }
                    }
                    dst[t * plane + (int64_t)y * W + x] += v;
                }
    }
}

} // namespace

extern "C" {

// Patch projection (see patch_project): mode ORC_F32 writes float, the other
// modes double.  Returns -1 on bad sizes (even filter volume, as Convolve
// asserts at src/Convolve.cpp:145).
int orc_patch_project(int mode, int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int K,
                      const float *filt, void *out) {
    if (W < 1 || H < 1 || F < 1 || C < 1 || K < 1 || fw < 1 || fh < 1 || ft < 1) return -1;
    if (((int64_t)fw * fh * ft) % 2 != 1 || mode < 0 || mode > 2) return -1;
    if (mode == ORC_F32) patch_project<float>(W, H, F, C, im, fw, fh, ft, K, filt, (float *)out);
    else patch_project<double>(W, H, F, C, im, fw, fh, ft, K, filt, (double *)out);
    return 0;
}

// NLMeans::apply after PatchPCA: the guide is the FP32 patch projection
// (the reference's arithmetic); then JointBilateral::apply(image, guide,
// sigma_s, sigma_s, INF, patch_sigma) with JointBilateral's AUTO choice
// (src/GaussTransform.cpp:553-564).  Returns -2 when AUTO picks EXACT or
// GKDTREE (not restated here), -1 on bad sizes.
int orc_nlmeans(int mode, int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int K,
                const float *filt, float spatial_sigma, float patch_sigma, void *out) {
    if (W < 1 || H < 1 || F < 1 || C < 1 || K < 1 || fw < 1 || fh < 1 || ft < 1) return -1;
    if (((int64_t)fw * fh * ft) % 2 != 1 || mode < 0 || mode > 2) return -1;
    float sx = spatial_sigma, sy = spatial_sigma, st = INFINITY;
    if (W == 1) sx = 0;
    if (H == 1) sy = 0;
    if (F == 1) st = 0;
    if ((W > 1 && sx == 0) || (H > 1 && sy == 0)) return -2;
    const bool fx = W > 1 && sx < 10 * W, fy = H > 1 && sy < 10 * H, ftm = F > 1 && st < 10 * F;
    const int pos = K + (fx ? 1 : 0) + (fy ? 1 : 0) + (ftm ? 1 : 0);
    const int szx = fx ? ((int)(sx * 6 + 1) | 1) : 1, szy = fy ? ((int)(sy * 6 + 1) | 1) : 1;
    const int szt = ftm ? ((int)(st * 6 + 1) | 1) : 1;
    if (szx * szy * szt < 16 || pos > 10) return -2;
    const bool grid = pos <= 4;
    const int64_t n = (int64_t)W * H * F;
    std::vector<float> guide((size_t)(K * n));
    patch_project<float>(W, H, F, C, im, fw, fh, ft, K, filt, guide.data());
    const int64_t ys = W, ts = (int64_t)W * H;
    View a{im, W, H, F, C, ys, ts, n}, b{guide.data(), W, H, F, K, ys, ts, n};
    if (mode == ORC_F32)
        return jb_rec<float>(mode, a, b, (float *)out, ys, ts, n, spatial_sigma, spatial_sigma, INFINITY,
                             patch_sigma, grid);
    return jb_rec<double>(mode, a, b, (double *)out, ys, ts, n, spatial_sigma, spatial_sigma, INFINITY,
                          patch_sigma, grid);
}

} // extern "C"
