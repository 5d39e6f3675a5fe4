// lp_hull.hip -- per-cluster monotone-chain convex hulls on gfx950.
//
// Reference: lidar_processing_lib/src/polygonizer.cpp:33-91 (Andrew's
// monotone chain over an (x, y)-sorted index list, popping while the
// double-precision cross product is <= 0, result resized to k-1) fed by the
// per-label gather of src/processor/src/processor.cpp:627-663 (cluster points
// as PointXY doubles in obstacle-index order; hull indices are positions in
// that list).
//
// With sign-exact orientation tests the chain returns the strictly convex
// hull vertices, counter-clockwise from the lexicographically smallest
// point (and [min, max] for collinear input).  That set is unchanged when
// points that are not vertices of a subset's hull are dropped, so hulls are
// built in rounds: every 256-point chunk of a cluster is sorted and chained
// by one wavefront, the chunk hull vertices become the next round's input,
// and a cluster whose input fits one chunk gets its final hull.  The
// orientation test evaluates the reference's expression in double and, when
// Shewchuk's error bound cannot certify its sign, falls back to an exact
// expansion sum (TwoProduct via FMA).
#include <algorithm>
#include <cfloat>

#include "lp_kernels.h"

namespace lp {

using namespace lpd;

// ---- exact orientation ---------------------------------------------------
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void two_diff(double a, double b, double& s, double& e) {
    s = a - b;
    const double bb = a - s;
    e = (a - (s + bb)) + (bb - b);
}
__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    e = fma(a, b, -p);
}

// Exact sign of (bx-ax)*(cy-ay) - (by-ay)*(cx-ax) when the double value
// `det` (sign ns) cannot be certified by the error bound.  Kept out of line:
// it is rare, and inlined at every call site its expansion arithmetic would
// crowd the instruction cache of the hot loops.
__device__ __attribute__((noinline)) int orient_tail(double ax, double ay, double bx, double by, double cx,
                                                     double cy, int ns) {
    double d[4][2];
    two_diff(bx, ax, d[0][0], d[0][1]);  // bx - ax
    two_diff(cy, ay, d[1][0], d[1][1]);  // cy - ay
    two_diff(by, ay, d[2][0], d[2][1]);  // by - ay
    two_diff(cx, ax, d[3][0], d[3][1]);  // cx - ax
    if (d[0][1] == 0.0 && d[1][1] == 0.0 && d[2][1] == 0.0 && d[3][1] == 0.0) {
        // Exact differences: det = (p1 + e1) - (p2 + e2) exactly, with the
        // product errors from FMA.  If both products are exact, the rounded
        // difference already has the exact sign (duplicates land here).
        double p1, e1, p2, e2;
        two_prod(d[0][0], d[1][0], p1, e1);
        two_prod(d[2][0], d[3][0], p2, e2);
        if (e1 == 0.0 && e2 == 0.0) return ns;
        // Four-term exact sum: grow-expansion over {p1, e1, -p2, -e2}.
        double t[4] = {e1, -e2, p1, -p2};
        double ex[4];
        int ne = 0;
        for (int k = 0; k < 4; ++k) {
            double q = t[k];
            for (int i = 0; i < ne; ++i) {
                double s, h;
                two_sum(q, ex[i], s, h);
                ex[i] = h;
                q = s;
            }
            ex[ne++] = q;
        }
        for (int i = ne - 1; i >= 0; --i) {
            if (ex[i] > 0.0) return 1;
            if (ex[i] < 0.0) return -1;
        }
        return 0;
    }
    double terms[16];
    int nt = 0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double p, e;
            two_prod(d[0][i], d[1][j], p, e);
            terms[nt++] = p;
            terms[nt++] = e;
            two_prod(d[2][i], d[3][j], p, e);
            terms[nt++] = -p;
            terms[nt++] = -e;
        }
    // Grow-expansion (Shewchuk 1997): exact nonoverlapping sum.
    double ex[17];
    int ne = 0;
    for (int t = 0; t < nt; ++t) {
        double q = terms[t];
        for (int i = 0; i < ne; ++i) {
            double s, h;
            two_sum(q, ex[i], s, h);
            ex[i] = h;
            q = s;
        }
        ex[ne++] = q;
    }
    for (int i = ne - 1; i >= 0; --i) {
        if (ex[i] > 0.0) return 1;
        if (ex[i] < 0.0) return -1;
    }
    return 0;
}

// Sign of (bx-ax)*(cy-ay) - (by-ay)*(cx-ax), exact.  *naive_sign receives
// the sign of the expression evaluated in double the way the reference does.
__device__ __forceinline__ int orient_exact(double ax, double ay, double bx, double by, double cx, double cy,
                                            int* naive_sign) {
    const double t1 = (bx - ax) * (cy - ay);
    const double t2 = (by - ay) * (cx - ax);
    const double det = t1 - t2;
    const int ns = (det > 0.0) - (det < 0.0);
    *naive_sign = ns;
    const double bound = 3.3306690738754716e-16 * (fabs(t1) + fabs(t2));
    if (det > bound || -det > bound) return ns;
    return orient_tail(ax, ay, bx, by, cx, cy, ns);
}

// ---- chunk hull ----------------------------------------------------------
// (x, y, idx) lexicographic order; idx breaks exact-duplicate ties.
__device__ __forceinline__ bool pt_less(const double* X, const double* Y, const int32_t* I, int a, int b) {
    if (X[a] != X[b]) return X[a] < X[b];
    if (Y[a] != Y[b]) return Y[a] < Y[b];
    return I[a] < I[b];
}

// One half of Andrew's chain (polygonizer.cpp:66-88) over the n sorted points
// P: the lower chain visits P[0..n), the upper chain P[n-1..0] starting from
// P[n-1], which the lower chain always ends with, so the two halves are
// independent and run on two lanes at once.  The top four stack points live
// in registers (a pop shifts them and refills the fourth from the stack, two
// pops before it is needed) and the points two ahead are loaded before the
// pop loop.  Pops while the exact orientation is <= 0.  Writes the chain to S
// and returns its length.
__device__ int half_chain(const double* X, const double* Y, const uint16_t* P, int n, bool upper, uint16_t* S,
                          uint32_t* inexact) {
    int kk = 0;
    double x1 = 0.0, y1 = 0.0, x2 = 0.0, y2 = 0.0, x3 = 0.0, y3 = 0.0, x4 = 0.0, y4 = 0.0;
    int ns;
    int p = upper ? P[n - 1] : P[0];
    double px = X[p], py = Y[p];
    int pn = (n > 1) ? (upper ? P[n - 2] : P[1]) : p;
    double pxn = X[pn], pyn = Y[pn];
    for (int t = 0; t < n; ++t) {
        const int pnn = (t + 2 < n) ? (upper ? P[n - 3 - t] : P[t + 2]) : pn;
        const double pxnn = X[pnn], pynn = Y[pnn];
        while (kk > 1) {
            const int o = orient_exact(x2, y2, x1, y1, px, py, &ns);
            *inexact += (o != ns) ? 1u : 0u;
            if (o > 0) break;
            --kk;
            x1 = x2;
            y1 = y2;
            x2 = x3;
            y2 = y3;
            x3 = x4;
            y3 = y4;
            if (kk > 3) {
                const int q = S[kk - 4];
                x4 = X[q];
                y4 = Y[q];
            }
        }
        S[kk++] = static_cast<uint16_t>(p);
        x4 = x3;
        y4 = y3;
        x3 = x2;
        y3 = y2;
        x2 = x1;
        y2 = y1;
        x1 = px;
        y1 = py;
        p = pn;
        px = pxn;
        py = pyn;
        pn = pnn;
        pxn = pxnn;
        pyn = pynn;
    }
    return kk;
}

// The same chain as half_chain, computed by the whole wavefront by divide
// and conquer: the sequence (L[0..n), reversed for the upper chain) is cut
// into 64 consecutive blocks of at most ceil(n / 64) points, each lane runs
// the chain over its block, and six levels merge neighbouring chains.  Two
// chains of consecutive blocks meet at their bridge (a, b): starting from the
// left chain's last and the right chain's first vertex, a is dropped while
// the turn (a - 1, a, b) is not strictly positive and b while (a, b, b + 1)
// is not, until both turns are; the merged chain is the left one up to a and
// the right one from b.  With sign-exact tests the chain of a sorted
// sequence is the unique strictly convex chain of its points, so the result
// equals the sequential walk's; the walk's depth (a pop loop per point) is
// replaced by six short merges, which pays on long, nearly convex halves
// (float-rounded straight edges keep most of their points).  S receives the
// chain (S has room for n entries); every lane returns its length.
__device__ int merge_half_chain(const double* X, const double* Y, const uint16_t* L, int n, bool upper, uint16_t* S,
                                uint32_t* inexact) {
    const int lane = lane_id();
    const int bs = (n + kWave - 1) / kWave;  // block size, at most 4 for a 256-point chunk
    for (int k = lane; k < n; k += kWave) S[k] = L[upper ? n - 1 - k : k];
    wave_sync();
    // Level 0: each lane's block, in place at its start.
    int len = 0;
    {
        const int b0 = lane * bs, b1 = min(b0 + bs, n);
        for (int k = b0; k < b1; ++k) {
            const int p = S[k];
            while (len > 1) {
                const int a = S[b0 + len - 2], b = S[b0 + len - 1];
                int ns;
                const int o = orient_exact(X[a], Y[a], X[b], Y[b], X[p], Y[p], &ns);
                *inexact += (o != ns) ? 1u : 0u;
                if (o > 0) break;
                --len;
            }
            S[b0 + len++] = static_cast<uint16_t>(p);
        }
    }
    wave_sync();
    // Merge levels: lane g (a multiple of 2w) joins its chain with lane
    // g + w's; the chain lengths travel by lane shuffles.
    for (int w = 1; w < kWave; w <<= 1) {
        const int right_len = __shfl(len, min(lane + w, kWave - 1), kWave);
        if ((lane & (2 * w - 1)) == 0 && lane + w < kWave && (lane + w) * bs < n) {
            const int lb = lane * bs, rb = (lane + w) * bs;
            const int nl = len, nr = right_len;
            int ia = nl - 1, ib = 0;
            bool moved = true;
            while (moved) {
                moved = false;
                while (ia > 0) {
                    const int a0 = S[lb + ia - 1], a = S[lb + ia], b = S[rb + ib];
                    int ns;
                    const int o = orient_exact(X[a0], Y[a0], X[a], Y[a], X[b], Y[b], &ns);
                    *inexact += (o != ns) ? 1u : 0u;
                    if (o > 0) break;
                    --ia;
                    moved = true;
                }
                while (ib < nr - 1) {
                    const int a = S[lb + ia], b = S[rb + ib], b1 = S[rb + ib + 1];
                    int ns;
                    const int o = orient_exact(X[a], Y[a], X[b], Y[b], X[b1], Y[b1], &ns);
                    *inexact += (o != ns) ? 1u : 0u;
                    if (o > 0) break;
                    ++ib;
                    moved = true;
                }
            }
            // Right part moves left to follow a (its destination never
            // passes its source).
            for (int k = ib; k < nr; ++k) S[lb + ia + 1 + k - ib] = S[rb + k];
            len = ia + 1 + nr - ib;
        }
        wave_sync();
    }
    return __shfl(len, 0, kWave);
}

// Splits the n >= 2 sorted points P by the line from P[0] to P[n-1]: LO gets
// P[0], the points strictly below it (in order), P[n-1]; UP the same with the
// points strictly above.  Only those can be vertices of the lower / upper
// chain, so each chain sees about half the points.  One wavefront; returns
// the two list lengths through *nlo, *nup.
__device__ void split_by_chord(const double* X, const double* Y, const uint16_t* P, int n, uint16_t* LO,
                               uint16_t* UP, int* nlo, int* nup) {
    const int lane = lane_id();
    const int a = P[0], b = P[n - 1];
    const double ax = X[a], ay = Y[a], bx = X[b], by = Y[b];
    int cl = 0, cu = 0;
    for (int i0 = 1; i0 < n - 1; i0 += kWave) {
        const int i = i0 + lane;
        int side = 0;
        int p = 0;
        if (i < n - 1) {
            int dummy;
            p = P[i];
            side = orient_exact(ax, ay, bx, by, X[p], Y[p], &dummy);
        }
        const uint64_t ml = __ballot(side < 0), mu = __ballot(side > 0);
        if (side < 0) LO[1 + cl + lanes_below(ml)] = static_cast<uint16_t>(p);
        if (side > 0) UP[1 + cu + lanes_below(mu)] = static_cast<uint16_t>(p);
        cl += static_cast<int>(__popcll(ml));
        cu += static_cast<int>(__popcll(mu));
    }
    if (lane == 0) {
        LO[0] = static_cast<uint16_t>(a);
        LO[cl + 1] = static_cast<uint16_t>(b);
        UP[0] = static_cast<uint16_t>(a);
        UP[cu + 1] = static_cast<uint16_t>(b);
    }
    *nlo = cl + 2;
    *nup = cu + 2;
}

// Parallel pruning of a half-chain list L[0..n) (sorted by (x, y); first and
// last kept): a point that does not make a strict turn of sign `turn` (+1
// lower, -1 upper) with two current neighbours of other coordinates lies on
// or beyond the segment between them, so it is no vertex of that half hull;
// all such points are dropped at once, and the pass repeats on the survivors.  The half hull of
// the survivors is the half hull of the list.  One wavefront; T is scratch of
// n entries.  Returns the new length.  *done becomes true when a pass drops
// nothing and keeps no point for a duplicate neighbour: every interior point
// then makes a strict turn, and an x-monotone chain of strict turns of one
// sign is its own half hull, so the chain walk can be skipped.  *inexact
// (per lane) counts tests whose double evaluation has another sign.
__device__ int prune_chain(const double* X, const double* Y, uint16_t* L, uint16_t* T, int n, int turn,
                           int passes, bool* done, uint32_t* inexact) {
    const int lane = lane_id();
    *done = n <= 2;
    for (int pass = 0; pass < passes && n > 2; ++pass) {
        int m = 0;
        bool dup_kept = false;
        for (int i0 = 0; i0 < n; i0 += kWave) {
            const int i = i0 + lane;
            bool keep = false;
            int b = 0;
            if (i < n) {
                b = L[i];
                keep = true;
                if (i > 0 && i < n - 1) {
                    const int a = L[i - 1], c = L[i + 1];
                    // A neighbour with the same coordinates is no witness:
                    // copies of one vertex would remove each other.
                    const bool dup = (X[a] == X[b] && Y[a] == Y[b]) || (X[c] == X[b] && Y[c] == Y[b]);
                    int ns;
                    const int o = orient_exact(X[a], Y[a], X[b], Y[b], X[c], Y[c], &ns);
                    *inexact += (o != ns) ? 1u : 0u;
                    keep = dup || o * turn > 0;
                    dup_kept = dup_kept || dup;
                }
            }
            const uint64_t mk = __ballot(keep);
            if (keep) T[m + lanes_below(mk)] = static_cast<uint16_t>(b);
            m += static_cast<int>(__popcll(mk));
        }
        wave_sync();
        if (m == n) {
            *done = __ballot(dup_kept) == 0;
            break;
        }
        for (int i = lane; i < m; i += kWave) L[i] = T[i];
        wave_sync();
        n = m;
    }
    return n;
}

constexpr int kPrunePasses = 6;
constexpr int kMergeChain = 32;  // half lists longer than this take the wavefront chain
constexpr int kBigPrunePasses = 32;
constexpr int kWavesPerBlock = 4;
constexpr int kHullRounds = 3;
constexpr int kHullBig = 4096;  // candidates one workgroup finishes after the rounds

// Andrew's monotone chain over the n points P[0..n) (sorted by (x, y, idx))
// with sign-exact orientation (polygonizer.cpp:66-90).  Writes the hull to
// S and returns its length; *inexact counts tests where the reference's
// double expression has a different sign than the exact one.
__device__ int chain_hull(const double* X, const double* Y, const int32_t* I, const uint16_t* P, int n, uint16_t* S,
                          uint32_t* inexact) {
    int kk = 0;
    if (n < 3) {
        // polygonizer.cpp:36-41: fewer than three points are returned as
        // they are (in obstacle-index order, the reference's list order).
        for (int i = 0; i < n; ++i) S[kk++] = P[i];
        for (int a = 1; a < kk; ++a)
            for (int b = a; b > 0 && I[S[b - 1]] > I[S[b]]; --b) {
                const uint16_t tmp = S[b];
                S[b] = S[b - 1];
                S[b - 1] = tmp;
            }
        return kk;
    }
    int ns;
    for (int i = 0; i < n; ++i) {
        const int p = P[i];
        while (kk > 1) {
            const int o = orient_exact(X[S[kk - 2]], Y[S[kk - 2]], X[S[kk - 1]], Y[S[kk - 1]], X[p], Y[p], &ns);
            *inexact += (o != ns) ? 1u : 0u;
            if (o > 0) break;
            --kk;
        }
        S[kk++] = static_cast<uint16_t>(p);
    }
    const int tt = kk + 1;
    for (int i = n - 2; i >= 0; --i) {
        const int p = P[i];
        while (kk >= tt) {
            const int o = orient_exact(X[S[kk - 2]], Y[S[kk - 2]], X[S[kk - 1]], Y[S[kk - 1]], X[p], Y[p], &ns);
            *inexact += (o != ns) ? 1u : 0u;
            if (o > 0) break;
            --kk;
        }
        S[kk++] = static_cast<uint16_t>(p);
    }
    return kk - 1;
}

// Order-preserving unsigned key of a float (-0 folded onto +0, so equal
// values give equal keys); NaN is excluded by the caller.
__device__ __forceinline__ uint32_t float_key(float v) {
    const uint32_t u = __float_as_uint(v + 0.0f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Bitonic sort of 64*S (key, value) pairs held in registers, element
// e = s * 64 + lane in slot s, by key alone.  Equal keys are the same (x, y)
// (or sort-last padding), and any value of such a group may survive an
// exchange: the index a vertex reports among exact duplicates is settled
// later by coordinates, so the network needs no tie rule.  Partners within a
// lane swap registers, others are exchanged with lane shuffles: no LDS round
// trips.
template <int S>
__device__ __forceinline__ void reg_bitonic(uint64_t (&key)[S], uint32_t (&val)[S]) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 2; k <= kWave * S; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kWave) {
                const int js = j / kWave;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (s & js) continue;
                    const int t = s | js;
                    const bool asc = ((s * kWave) & k) == 0;
                    const bool gt = key[s] > key[t];
                    if (gt == asc) {
                        const uint64_t tk = key[s];
                        key[s] = key[t];
                        key[t] = tk;
                        const uint32_t tv = val[s];
                        val[s] = val[t];
                        val[t] = tv;
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    // Partners inside a 16-lane row move by DPP (lane_xor),
                    // the others through a shuffle.
                    uint64_t ok;
                    uint32_t ov;
                    if (j < 16) {
                        const uint32_t klo = lane_xor(static_cast<uint32_t>(key[s]), j);
                        const uint32_t khi = lane_xor(static_cast<uint32_t>(key[s] >> 32), j);
                        ok = (static_cast<uint64_t>(khi) << 32) | klo;
                        ov = lane_xor(val[s], j);
                    } else {
                        ok = __shfl_xor(key[s], j, kWave);
                        ov = __shfl_xor(val[s], j, kWave);
                    }
                    const bool asc = ((s * kWave + lane) & k) == 0;
                    const bool lower = (lane & j) == 0;
                    const bool other_less = ok < key[s];
                    if (other_less == (lower == asc)) {
                        key[s] = ok;
                        val[s] = ov;
                    }
                }
            }
        }
    }
}

// The whole chunk step for float coordinates, in registers: 64-bit keys of
// the n loaded points, the Akl-Toussaint cull against the distinct extremes
// (key minima / maxima: x-major gives the leftmost and rightmost points,
// y-major the lowest and highest; any extreme points serve), then the sort,
// the sort by (x, y), the duplicate drop (unless fewer than three points
// survive the cull) and the write of X, Y, I in sorted order with P the
// identity.  Returns the new count; *ns_in receives the number of points the
// cull kept.
template <int S>
__device__ int sort_dedupe_chunk(double* X, double* Y, int32_t* I, uint16_t* P, uint64_t (&key)[S],
                                 uint32_t (&val)[S], int ns);

// The ns real keys (key != ~0) of S slots moved into the first of S2 < S
// slots through LDS (X holds the keys, P the values: neither is read again
// before the sorted write), then sorted there.
template <int S, int S2>
__device__ int compact_sort_chunk(double* X, double* Y, int32_t* I, uint16_t* P, uint64_t (&key)[S],
                                  uint32_t (&val)[S], int ns) {
    const int lane = lane_id();
    uint64_t* ks = reinterpret_cast<uint64_t*>(X);
    int before = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const bool real = key[s] != ~0ull;
        const uint64_t m = __ballot(real);
        if (real) {
            const int q = before + static_cast<int>(lanes_below(m));
            ks[q] = key[s];
            P[q] = static_cast<uint16_t>(val[s]);
        }
        before += static_cast<int>(__popcll(m));
    }
    wave_sync();
    uint64_t key2[S2];
    uint32_t val2[S2];
#pragma unroll
    for (int s = 0; s < S2; ++s) {
        const int e = s * kWave + lane;
        key2[s] = e < ns ? ks[e] : ~0ull;
        val2[s] = e < ns ? static_cast<uint32_t>(P[e]) : static_cast<uint32_t>(kHullChunk + e);
    }
    wave_sync();
    return sort_dedupe_chunk<S2>(X, Y, I, P, key2, val2, ns);
}

template <int S>
__device__ int cull_sort_float_chunk(double* X, double* Y, int32_t* I, uint16_t* P, int n, int* ns_in) {
    const int lane = lane_id();
    uint64_t key[S];
    uint32_t val[S];
    uint64_t mnx = ~0ull, mxx = 0ull, mny = ~0ull, mxy = 0ull;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int e = s * kWave + lane;
        if (e < n) {
            key[s] = (static_cast<uint64_t>(float_key(static_cast<float>(X[e]))) << 32) |
                     float_key(static_cast<float>(Y[e]));
            val[s] = static_cast<uint32_t>(e);
            const uint64_t ky = (key[s] << 32) | (key[s] >> 32);
            mnx = min(mnx, key[s]);
            mxx = max(mxx, key[s]);
            mny = min(mny, ky);
            mxy = max(mxy, ky);
        } else {
            key[s] = ~0ull;
            val[s] = static_cast<uint32_t>(kHullChunk + e);
        }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        mnx = min(mnx, static_cast<uint64_t>(__shfl_xor(mnx, o, kWave)));
        mxx = max(mxx, static_cast<uint64_t>(__shfl_xor(mxx, o, kWave)));
        mny = min(mny, static_cast<uint64_t>(__shfl_xor(mny, o, kWave)));
        mxy = max(mxy, static_cast<uint64_t>(__shfl_xor(mxy, o, kWave)));
    }
    // Counter-clockwise: leftmost, lowest, rightmost, highest; repeated
    // points dropped so no edge has zero length.
    const uint64_t ext[4] = {mnx, (mny << 32) | (mny >> 32), mxx, (mxy << 32) | (mxy >> 32)};
    // Appended by static-index selects, so the four corners stay in
    // registers (a dynamically indexed array would live in scratch memory).
    double px[4] = {0.0, 0.0, 0.0, 0.0}, py[4] = {0.0, 0.0, 0.0, 0.0};
    int npoly = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool fresh = npoly == 0 || ext[e] != ext[e > 0 ? e - 1 : 0];
        if (fresh) {
            const double ex_x = static_cast<double>(key_float(static_cast<uint32_t>(ext[e] >> 32)));
            const double ex_y = static_cast<double>(key_float(static_cast<uint32_t>(ext[e])));
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (t == npoly) {
                    px[t] = ex_x;
                    py[t] = ex_y;
                }
            }
            ++npoly;
        }
    }
    if (npoly > 1 && ext[3] == ext[0]) --npoly;
    int ns = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int e = s * kWave + lane;
        bool keep = e < n;
        if (keep && npoly >= 3) {
            const double qx = static_cast<double>(key_float(static_cast<uint32_t>(key[s] >> 32)));
            const double qy = static_cast<double>(key_float(static_cast<uint32_t>(key[s])));
            bool inside = true;
            int dummy;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k < npoly && inside) {
                    const bool wrap_to_0 = k + 1 >= npoly;
                    const double bx = wrap_to_0 ? px[0] : px[k < 3 ? k + 1 : 0];
                    const double by = wrap_to_0 ? py[0] : py[k < 3 ? k + 1 : 0];
                    inside = orient_exact(px[k], py[k], bx, by, qx, qy, &dummy) > 0;
                }
            }
            keep = !inside;
        }
        if (e < n && !keep) key[s] = ~0ull;  // culled: sorts last
        ns += static_cast<int>(__popcll(__ballot(keep)));
    }
    *ns_in = ns;
    // Fewer survivors than slots: sort them in fewer slots (the network's
    // cost grows with the slot count).
    if constexpr (S == 4) {
        if (ns <= kWave) return compact_sort_chunk<4, 1>(X, Y, I, P, key, val, ns);
        if (ns <= 2 * kWave) return compact_sort_chunk<4, 2>(X, Y, I, P, key, val, ns);
    }
    if constexpr (S == 2) {
        if (ns <= kWave) return compact_sort_chunk<2, 1>(X, Y, I, P, key, val, ns);
    }
    return sort_dedupe_chunk<S>(X, Y, I, P, key, val, ns);
}

// Sort of the (key, value) slots (culled and padding keys are ~0 and sort
// last), the duplicate drop and the write of X, Y, I in sorted order with P
// the identity; ns is the number of real keys.
template <int S>
__device__ int sort_dedupe_chunk(double* X, double* Y, int32_t* I, uint16_t* P, uint64_t (&key)[S],
                                 uint32_t (&val)[S], int ns) {
    const int lane = lane_id();
    const bool keep_all = ns < 3;
    reg_bitonic<S>(key, val);
    int32_t idx[S];
    uint64_t prev[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int e = s * kWave + lane;
        idx[s] = e < ns ? I[val[s]] : 0;
        const uint64_t up = __shfl_up(key[s], 1, kWave);
        const uint64_t wrap = __shfl(key[s > 0 ? s - 1 : 0], kWave - 1, kWave);
        prev[s] = lane > 0 ? up : wrap;
    }
    int nd = 0;
    uint64_t kmask[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int e = s * kWave + lane;
        kmask[s] = __ballot(e < ns && (keep_all || e == 0 || key[s] != prev[s]));
        nd += static_cast<int>(__popcll(kmask[s]));
    }
    if (!keep_all && nd == 1 && ns >= 2) {
        kmask[0] |= 2ull;  // one repeated point: keep two copies
        nd = 2;
    }
    wave_sync();  // every lane has read I[] before it is rewritten
    int before = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if ((kmask[s] >> lane) & 1ull) {
            const int q = before + static_cast<int>(lanes_below(kmask[s]));
            X[q] = static_cast<double>(key_float(static_cast<uint32_t>(key[s] >> 32)));
            Y[q] = static_cast<double>(key_float(static_cast<uint32_t>(key[s])));
            I[q] = idx[s];
            P[q] = static_cast<uint16_t>(q);
        }
        before += static_cast<int>(__popcll(kmask[s]));
    }
    wave_sync();
    return nd;
}

// Chunk plan for round r, one 1024-thread workgroup per frame walking the
// clusters in tiles with a running carry: a cluster that is not final yet
// needs ceil(cnt / 256) chunks (one if it fits: that chunk's hull is final);
// hc_choff receives the exclusive scan of the chunk counts, hc_chunk_c /
// hc_chunk_i / hc_chunk_b each chunk's cluster, index, size and the cluster's
// first candidate, and the frame's total goes to kFsHullChunks.  Round 0 also clears the hull count and flag,
// and every round clears the next round's candidate counts.
constexpr int kPlanThreads = 1024;
constexpr int kPlanItems = 4;

__global__ __launch_bounds__(kPlanThreads) void k_hull_plan(DevBuf d, Layout L, int r) {
    __shared__ uint32_t lds[kPlanThreads / kWave + 1];
    const int f = blockIdx.x;
    const int K = static_cast<int>(d.fs[f * kFsCount + kFsClusters]);
    const size_t ob = (size_t)f * L.N;
    const size_t BN = (size_t)L.B * L.N;
    uint32_t carry = 0;
    for (int base = 0; base < K; base += kPlanThreads * kPlanItems) {
        uint32_t v[kPlanItems];
        uint32_t local = 0;
        const int b = base + static_cast<int>(threadIdx.x) * kPlanItems;
        uint32_t cn[kPlanItems];
#pragma unroll
        for (int q = 0; q < kPlanItems; ++q) {
            const int c = b + q;
            v[q] = 0;
            cn[q] = 0;
            if (c < K) {
                if (r == 0) {
                    d.hull_count[ob + c] = 0;
                    d.hull_flag[ob + c] = 0;
                }
                if (r == 0 || d.hull_count[ob + c] == 0) {
                    const uint32_t cnt = d.hc_count[(r & 1) * BN + ob + c];
                    v[q] = cnt <= static_cast<uint32_t>(kHullChunk) ? 1u : (cnt + kHullChunk - 1) / kHullChunk;
                    cn[q] = cnt;
                }
                d.hc_count[((r + 1) & 1) * BN + ob + c] = 0;
                local += v[q];
            }
        }
        uint32_t total;
        uint32_t ex = block_exclusive_sum<kPlanThreads>(local, lds, &total) + carry;
#pragma unroll
        for (int q = 0; q < kPlanItems; ++q) {
            if (b + q < K) {
                d.hc_choff[ob + b + q] = ex;
                // The chunk map of each of the cluster's chunks, so a round's
                // wavefront has its chunk's cluster, place and size with one
                // round of loads.  A round has at most N chunks (a cluster's
                // chunk count never exceeds its candidate count).
                if (v[q]) {
                    const uint32_t base = d.cl_off[ob + b + q];
                    const uint32_t fin = cn[q] <= static_cast<uint32_t>(kHullChunk) ? 0x80000000u : 0u;
                    for (uint32_t k = 0; k < v[q] && ex + k < static_cast<uint32_t>(L.N); ++k) {
                        const uint32_t nk = min(cn[q] - k * kHullChunk, static_cast<uint32_t>(kHullChunk));
                        d.hc_chunk_c[ob + ex + k] = static_cast<int32_t>(static_cast<uint32_t>(b + q) | fin);
                        d.hc_chunk_i[ob + ex + k] = k | nk << 16;
                        d.hc_chunk_b[ob + ex + k] = base;
                    }
                }
            }
            ex += v[q];
        }
        carry += total;
    }
    if (threadIdx.x == 0) d.fs[f * kFsCount + kFsHullChunks] = carry;
}

// Register budget for three wavefronts per SIMD (168 VGPRs, a few spills in
// the cold paths): the chunks' sort and prune phases are VALU- and
// LDS-latency-bound, and a third wavefront hides more of it than two.
__global__ __launch_bounds__(256, 3) void k_hull_round(DevBuf d, Layout L, int r) {
    __shared__ double sx[kWavesPerBlock][kHullChunk];
    __shared__ double sy[kWavesPerBlock][kHullChunk];
    __shared__ int32_t si[kWavesPerBlock][kHullChunk];
    __shared__ uint16_t perm[kWavesPerBlock][kHullChunk];
    __shared__ uint16_t stk[kWavesPerBlock][2 * kHullChunk];
    __shared__ uint16_t lol[kWavesPerBlock][kHullChunk];
    __shared__ uint16_t upl[kWavesPerBlock][kHullChunk];
    __shared__ int nlo[kWavesPerBlock];
    __shared__ int nup[kWavesPerBlock];
    int tile, f;
    xcd_tile_frame(tile, f);
    const uint32_t total = min(d.fs[f * kFsCount + kFsHullChunks], static_cast<uint32_t>(L.N));
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    const size_t ob = (size_t)f * L.N;
    const size_t BN = (size_t)L.B * L.N;
    double* X = sx[w];
    double* Y = sy[w];
    int32_t* I = si[w];
    uint16_t* P = perm[w];
    uint16_t* S = stk[w];
    const size_t in_b = (size_t)(r & 1) * BN;
    const size_t out_b = (size_t)((r + 1) & 1) * BN;
    for (uint32_t t = tile * kWavesPerBlock + w; t < total; t += gridDim.x * kWavesPerBlock) {
        const uint32_t cw = static_cast<uint32_t>(d.hc_chunk_c[ob + t]);
        const uint32_t iw = d.hc_chunk_i[ob + t];
        const uint32_t base = d.hc_chunk_b[ob + t];
        const int c = static_cast<int>(cw & 0x7fffffffu);
        const bool final_chunk = (cw >> 31) != 0;
        const uint32_t lo = (iw & 0xffffu) * kHullChunk;
        const int n = static_cast<int>(iw >> 16);
        for (int i = lane; i < n; i += kWave) {
            const size_t src = in_b + ob + base + lo + i;
            X[i] = d.hcx[src];
            Y[i] = d.hcy[src];
            I[i] = d.hci[src];
        }
        wave_sync();
        // Coordinates that are all floats (every obstacle point) take the
        // register path: keys, cull, sort and duplicate drop in one pass.
        // Other lists use the LDS forms below.
        bool as_float = true;
        for (int i = lane; i < n; i += kWave) {
            const double x = X[i], y = Y[i];
            as_float = as_float && static_cast<double>(static_cast<float>(x)) == x &&
                       static_cast<double>(static_cast<float>(y)) == y;
        }
        as_float = __ballot(!as_float) == 0;
        int ns = 0;
        int ns_in = 0;
        // Exact duplicates (same x and y) are adjacent after the sort.  A
        // second copy of a point is never a hull vertex of a set with two or
        // more distinct points (the reference's chain pops it: its turn is
        // zero), so only the first copy stays; a list of one repeated point
        // keeps two copies, the reference's result for it.  Vertical surfaces
        // repeat one (x, y) once per beam, so this often shrinks a chunk
        // several times over.  Fewer than three points are kept as they are.
        if (as_float) {
            if (n <= kWave)
                ns = cull_sort_float_chunk<1>(X, Y, I, P, n, &ns_in);
            else if (n <= 2 * kWave)
                ns = cull_sort_float_chunk<2>(X, Y, I, P, n, &ns_in);
            else
                ns = cull_sort_float_chunk<4>(X, Y, I, P, n, &ns_in);
        } else {
        // Akl-Toussaint cull: points strictly inside the quadrilateral of the
        // leftmost, lowest, rightmost and highest points (counter-clockwise
        // along the hull boundary) are no hull vertices; exact tests only.
        int ex[4] = {0, 0, 0, 0};
        for (int i = lane; i < n; i += kWave) {
            if (X[i] < X[ex[0]] || (X[i] == X[ex[0]] && i < ex[0])) ex[0] = i;
            if (Y[i] < Y[ex[1]] || (Y[i] == Y[ex[1]] && i < ex[1])) ex[1] = i;
            if (X[i] > X[ex[2]] || (X[i] == X[ex[2]] && i < ex[2])) ex[2] = i;
            if (Y[i] > Y[ex[3]] || (Y[i] == Y[ex[3]] && i < ex[3])) ex[3] = i;
        }
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            int c0 = __shfl_xor(ex[0], o, kWave), c1 = __shfl_xor(ex[1], o, kWave);
            int c2 = __shfl_xor(ex[2], o, kWave), c3 = __shfl_xor(ex[3], o, kWave);
            if (X[c0] < X[ex[0]] || (X[c0] == X[ex[0]] && c0 < ex[0])) ex[0] = c0;
            if (Y[c1] < Y[ex[1]] || (Y[c1] == Y[ex[1]] && c1 < ex[1])) ex[1] = c1;
            if (X[c2] > X[ex[2]] || (X[c2] == X[ex[2]] && c2 < ex[2])) ex[2] = c2;
            if (Y[c3] > Y[ex[3]] || (Y[c3] == Y[ex[3]] && c3 < ex[3])) ex[3] = c3;
        }
        // The distinct extremes, in counter-clockwise order; repeated ones
        // (e.g. the leftmost point is also the lowest) are dropped so no edge
        // has zero length.  With fewer than three there is no interior.
        int poly[4];
        int npoly = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int v = ex[e];
            if (npoly == 0 || X[v] != X[poly[npoly - 1]] || Y[v] != Y[poly[npoly - 1]]) poly[npoly++] = v;
        }
        if (npoly > 1 && X[poly[npoly - 1]] == X[poly[0]] && Y[poly[npoly - 1]] == Y[poly[0]]) --npoly;
        for (int i0 = 0; i0 < n; i0 += kWave) {
            const int i = i0 + lane;
            bool keep = false;
            if (i < n) {
                int dummy;
                keep = true;
                if (npoly >= 3) {
                    bool inside = true;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        if (e < npoly && inside) {
                            const int a = poly[e], b = poly[e + 1 < npoly ? e + 1 : 0];
                            inside = orient_exact(X[a], Y[a], X[b], Y[b], X[i], Y[i], &dummy) > 0;
                        }
                    }
                    keep = !inside;
                }
            }
            const uint64_t mask = __ballot(keep);
            if (keep) P[ns + lanes_below(mask)] = static_cast<uint16_t>(i);
            ns += static_cast<int>(__popcll(mask));
        }
        ns_in = ns;
        int m = 1;
        while (m < ns) m <<= 1;
        for (int i = ns + lane; i < m; i += kWave) P[i] = static_cast<uint16_t>(n + i);  // padding sorts last
        wave_sync();
        // Bitonic sort of the permutation; padding entries (>= n) sort last.
        for (int k = 2; k <= m; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int q = lane; q < (m >> 1); q += kWave) {
                    const int a = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                    const int b = a + j;
                    const int pa = P[a], pb = P[b];
                    const bool gt = (pa >= n) ? (pb < n || pa > pb) : (pb < n && pt_less(X, Y, I, pb, pa));
                    const bool up = (a & k) == 0;
                    if (gt == up) {
                        P[a] = static_cast<uint16_t>(pb);
                        P[b] = static_cast<uint16_t>(pa);
                    }
                }
                wave_sync();
            }
        }
        if (ns >= 3) {
            int nd = 0;
            for (int i0 = 0; i0 < ns; i0 += kWave) {
                const int i = i0 + lane;
                bool keep = false;
                int p = 0;
                if (i < ns) {
                    p = P[i];
                    const int q = P[i > 0 ? i - 1 : 0];
                    keep = i == 0 || X[p] != X[q] || Y[p] != Y[q];
                }
                const uint64_t mk = __ballot(keep);
                if (keep) S[nd + lanes_below(mk)] = static_cast<uint16_t>(p);
                nd += static_cast<int>(__popcll(mk));
            }
            if (nd == 1) {
                if (lane == 0) S[1] = P[1];
                nd = 2;
            }
            wave_sync();
            for (int i = lane; i < nd; i += kWave) P[i] = S[i];
            wave_sync();
            ns = nd;
        }
        }
        // Lane 0 the lower chain, lane 1 the upper one (in S + kHullChunk);
        // vertex i of the hull is lower[i] for i < kl, then upper[i - kl + 1].
        // Fewer than three input points take the reference's short path; two
        // distinct points left after the duplicates go through the chain,
        // which returns them in (x, y) order.
        if (ns_in < 3) {
            if (lane == 0) {
                uint32_t inexact = 0;
                nlo[w] = chain_hull(X, Y, I, P, ns, S, &inexact);
                nup[w] = 2;
            }
        } else {
            int cl, cu;
            split_by_chord(X, Y, P, ns, lol[w], upl[w], &cl, &cu);
            wave_sync();
            bool dl, du;
            uint32_t inexact = 0;
            cl = prune_chain(X, Y, lol[w], S, cl, 1, kPrunePasses, &dl, &inexact);
            cu = prune_chain(X, Y, upl[w], S + kHullChunk, cu, -1, kPrunePasses, &du, &inexact);
            if (dl && du) {
                // Both halves are already chains: lower as listed, upper
                // from the right end back.
                for (int i = lane; i < cl; i += kWave) S[i] = lol[w][i];
                for (int i = lane; i < cu; i += kWave) S[kHullChunk + i] = upl[w][cu - 1 - i];
                if (lane == 0) {
                    nlo[w] = cl;
                    nup[w] = cu;
                }
            } else if (max(cl, cu) > kMergeChain) {
                // Long halves: the wavefront's divide-and-conquer chain.
                const int kl2 = merge_half_chain(X, Y, lol[w], cl, false, S, &inexact);
                const int ku2 = merge_half_chain(X, Y, upl[w], cu, true, S + kHullChunk, &inexact);
                if (lane == 0) {
                    nlo[w] = kl2;
                    nup[w] = ku2;
                }
            } else if (lane < 2) {
                // One call with per-lane arguments, so the two chains run in
                // lock-step instead of one after the other.
                const int len = half_chain(X, Y, lane == 0 ? lol[w] : upl[w], lane == 0 ? cl : cu, lane == 1,
                                           lane == 1 ? S + kHullChunk : S, &inexact);
                if (lane == 0)
                    nlo[w] = len;
                else
                    nup[w] = len;
            }
            inexact = wave_sum(inexact);
            if (lane == 0 && inexact) {
                atomicAdd(&d.fs[f * kFsCount + kFsHullInexact], inexact);
                atomicOr(&d.hull_flag[ob + c], 1u);
            }
        }
        wave_sync();
        const int kl = nlo[w];
        const int kk = kl + nup[w] - 2;
        auto vert = [&](int i) -> int { return i < kl ? S[i] : S[kHullChunk + i - kl + 1]; };
        if (final_chunk) {
            for (int i = lane; i < kk; i += kWave) d.hull_idx[ob + base + i] = I[vert(i)];
            if (lane == 0) d.hull_count[ob + c] = static_cast<uint32_t>(kk);
        } else {
            uint32_t pos = 0;
            if (lane == 0) pos = atomicAdd(&d.hc_count[out_b + ob + c], static_cast<uint32_t>(kk));
            pos = __shfl(pos, 0, kWave);
            for (int i = lane; i < kk; i += kWave) {
                const size_t dst = out_b + ob + base + pos + i;
                const int v = vert(i);
                d.hcx[dst] = X[v];
                d.hcy[dst] = Y[v];
                d.hci[dst] = I[v];
            }
        }
        wave_sync();
    }
}

// Clusters whose candidates still exceed one chunk after the rounds (hulls
// with hundreds of vertices): one workgroup sorts up to kHullBig candidates
// in LDS and chains them.
__global__ __launch_bounds__(1024) void k_hull_big(DevBuf d, Layout L, int r_in) {
    __shared__ double X[kHullBig];
    __shared__ double Y[kHullBig];
    __shared__ int32_t I[kHullBig];
    __shared__ uint16_t P[kHullBig];
    __shared__ uint16_t S[2 * kHullBig];
    __shared__ uint16_t LO[kHullBig];
    __shared__ uint16_t UP[kHullBig];
    __shared__ int nlo, nup, nlist[2];
    __shared__ bool nlist_done;
    __shared__ int open_list[1024];
    __shared__ int open_n;
    const int f = blockIdx.y;
    const int K = static_cast<int>(d.fs[f * kFsCount + kFsClusters]);
    const int tid = threadIdx.x;
    const size_t ob = (size_t)f * L.N;
    const size_t BN = (size_t)L.B * L.N;
    const size_t in_b = (size_t)(r_in & 1) * BN;
    // Workgroup x takes the clusters c = x (mod gridDim.x): its threads test
    // 1024 of them at once and list those without a hull yet.
    const int T = static_cast<int>(gridDim.x);
    for (int cb = 0; cb < K; cb += 1024 * T) {
        if (tid == 0) open_n = 0;
        __syncthreads();
        const int ct = cb + tid * T + static_cast<int>(blockIdx.x);
        if (ct < K && d.hull_count[ob + ct] == 0) open_list[atomicAdd(&open_n, 1)] = ct;
        __syncthreads();
        const int nopen = open_n;
        for (int jo = 0; jo < nopen; ++jo) {
            const int c = open_list[jo];
            const uint32_t cnt = d.hc_count[in_b + ob + c];
            if (cnt > static_cast<uint32_t>(kHullBig)) {
                if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrHullCapacity);
                continue;
            }
            const int n = static_cast<int>(cnt);
            const uint32_t base = d.cl_off[ob + c];
            int m = 1;
            while (m < n) m <<= 1;
            for (int i = tid; i < m; i += 1024) {
                if (i < n) {
                    const size_t src = in_b + ob + base + i;
                    X[i] = d.hcx[src];
                    Y[i] = d.hcy[src];
                    I[i] = d.hci[src];
                }
                P[i] = static_cast<uint16_t>(i);
            }
            __syncthreads();
            for (int k = 2; k <= m; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int q = tid; q < (m >> 1); q += 1024) {
                        const int a = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                        const int b = a + j;
                        const int pa = P[a], pb = P[b];
                        const bool gt = (pa >= n) ? (pb < n || pa > pb) : (pb < n && pt_less(X, Y, I, pb, pa));
                        if (gt == ((a & k) == 0)) {
                            P[a] = static_cast<uint16_t>(pb);
                            P[b] = static_cast<uint16_t>(pa);
                        }
                    }
                    __syncthreads();
                }
            }
            if (n < 3) {
                if (tid == 0) {
                    uint32_t inexact = 0;
                    nlo = chain_hull(X, Y, I, P, n, S, &inexact);
                    nup = 2;
                }
            } else {
                if (tid < kWave) {
                    // Exact duplicates (copies from different chunks) are
                    // adjacent now: keep the first of each, or two copies of a
                    // single repeated point (as in the chunk rounds).
                    const int lane = tid;
                    int nd = 0;
                    for (int i0 = 0; i0 < n; i0 += kWave) {
                        const int i = i0 + lane;
                        bool keep = false;
                        int p = 0;
                        if (i < n) {
                            p = P[i];
                            const int q = P[i > 0 ? i - 1 : 0];
                            keep = i == 0 || X[p] != X[q] || Y[p] != Y[q];
                        }
                        const uint64_t mk = __ballot(keep);
                        if (keep) LO[nd + lanes_below(mk)] = static_cast<uint16_t>(p);
                        nd += static_cast<int>(__popcll(mk));
                    }
                    if (nd == 1) {
                        if (lane == 0) LO[1] = P[1];
                        nd = 2;
                    }
                    wave_sync();
                    for (int i = lane; i < nd; i += kWave) P[i] = LO[i];
                    wave_sync();
                    int cl, cu;
                    split_by_chord(X, Y, P, nd, LO, UP, &cl, &cu);
                    wave_sync();
                    bool dl, du;
                    uint32_t inexact = 0;
                    cl = prune_chain(X, Y, LO, S, cl, 1, kBigPrunePasses, &dl, &inexact);
                    cu = prune_chain(X, Y, UP, S + kHullBig, cu, -1, kBigPrunePasses, &du, &inexact);
                    if (dl && du) {
                        for (int i = tid; i < cl; i += kWave) S[i] = LO[i];
                        for (int i = tid; i < cu; i += kWave) S[kHullBig + i] = UP[cu - 1 - i];
                    }
                    inexact = wave_sum(inexact);
                    if (tid == 0) {
                        nlist[0] = cl;
                        nlist[1] = cu;
                        nlist_done = dl && du;
                        if (dl && du) {
                            nlo = cl;
                            nup = cu;
                        }
                        if (inexact) {
                            atomicAdd(&d.fs[f * kFsCount + kFsHullInexact], inexact);
                            atomicOr(&d.hull_flag[ob + c], 1u);
                        }
                    }
                }
                __syncthreads();
                if (!nlist_done && tid < 2 * kWave) {
                    // Wavefront 0 the lower chain, wavefront 1 the upper one, each
                    // by the divide-and-conquer chain.
                    const int wv = tid / kWave;
                    uint32_t inexact = 0;
                    const int len = merge_half_chain(X, Y, wv == 0 ? LO : UP, nlist[wv], wv == 1,
                                                     wv == 1 ? S + kHullBig : S, &inexact);
                    inexact = wave_sum(inexact);
                    if (lane_id() == 0) {
                        if (wv == 0)
                            nlo = len;
                        else
                            nup = len;
                        if (inexact) {
                            atomicAdd(&d.fs[f * kFsCount + kFsHullInexact], inexact);
                            atomicOr(&d.hull_flag[ob + c], 1u);
                        }
                    }
                }
            }
            __syncthreads();
            const int kl = nlo;
            const int kk = kl + nup - 2;
            for (int i = tid; i < kk; i += 1024)
                d.hull_idx[ob + base + i] = I[i < kl ? S[i] : S[kHullBig + i - kl + 1]];
            if (tid == 0) d.hull_count[ob + c] = static_cast<uint32_t>(kk);
            __syncthreads();
        }
        __syncthreads();
    }
}

// Coordinates of member m of the cluster at `base`: the obstacle point, or
// for standalone lists the staged (x, y) pair.
__device__ __forceinline__ void member_xy(const DevBuf& d, const double2* list_xy, size_t ob, uint32_t base, int32_t m,
                                          double& x, double& y) {
    if (list_xy) {
        x = list_xy[base + m].x;
        y = list_xy[base + m].y;
    } else {
        x = d.ox[ob + m];
        y = d.oy[ob + m];
    }
}

// Clusters where some orientation test's double value had another sign than
// the exact one: the chunked, culled hull above is the exact hull, which the
// reference's rounded predicate need not produce.  Such a cluster is redone
// as polygonizer.cpp:33-91 computes it: all its points sorted by (x, y), then
// Andrew's chain popping while the double expression
// (p2.x-p1.x)*(p3.y-p1.y) - (p2.y-p1.y)*(p3.x-p1.x) <= 0, evaluated without
// contraction.  One workgroup per flagged cluster; the points are sorted in
// the first candidate buffer (free after k_hull_big) by an ascending-only
// bitonic network whose padding positions never move, and the lower and
// upper chains run on lanes 0 and 1 with their stacks in the second buffer.
// Exact duplicates may be ordered either way (their coordinates agree);
// distinct x closer than DBL_EPSILON, where the reference's comparator
// treats x as equal and is not an ordering, are reported as an error.
__device__ __forceinline__ bool ref_ccw_or_collinear(double ax, double ay, double bx, double by, double cx,
                                                     double cy) {
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax) <= 0;
}

__device__ __forceinline__ bool rec_less(const double* X, const double* Y, const int32_t* I, int a, int b) {
    if (X[a] != X[b]) return X[a] < X[b];
    if (Y[a] != Y[b]) return Y[a] < Y[b];
    return I[a] < I[b];
}

__global__ __launch_bounds__(256) void k_hull_ref(DevBuf d, Layout L, const double2* list_xy) {
    __shared__ int nlo, nup;
    int tile, f;
    xcd_tile_frame(tile, f);
    const int K = static_cast<int>(d.fs[f * kFsCount + kFsClusters]);
    const int tid = threadIdx.x;
    const size_t ob = (size_t)f * L.N;
    const size_t BN = (size_t)L.B * L.N;
    for (int c = tile; c < K; c += gridDim.x) {
        if (d.hull_flag[ob + c] == 0) continue;
        const int n = static_cast<int>(d.cl_size[ob + c]);
        if (n < 3) continue;
        const uint32_t base = d.cl_off[ob + c];
        double* X = d.hcx + ob + base;
        double* Y = d.hcy + ob + base;
        int32_t* I = d.hci + ob + base;
        const int32_t* members = d.cl_members + ob + base;
        for (int i = tid; i < n; i += blockDim.x) {
            const int32_t m = members[i];
            double x, y;
            member_xy(d, list_xy, ob, base, m, x, y);
            X[i] = x;
            Y[i] = y;
            I[i] = m;
        }
        __syncthreads();
        int m2 = 1;
        while (m2 < n) m2 <<= 1;
        for (int k = 2; k <= m2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int q = tid; q < (m2 >> 1); q += blockDim.x) {
                    int a, b;
                    if (j == (k >> 1)) {
                        // First step of a merge: mirror pairs inside each block of k.
                        const int blk = q / j, t = q % j;
                        a = blk * k + t;
                        b = blk * k + k - 1 - t;
                    } else {
                        a = ((q & ~(j - 1)) << 1) | (q & (j - 1));
                        b = a + j;
                    }
                    if (b < n && rec_less(X, Y, I, b, a)) {
                        const double tx = X[a], ty = Y[a];
                        const int32_t ti = I[a];
                        X[a] = X[b];
                        Y[a] = Y[b];
                        I[a] = I[b];
                        X[b] = tx;
                        Y[b] = ty;
                        I[b] = ti;
                    }
                }
                __syncthreads();
            }
        }
        for (int i = 1 + tid; i < n; i += blockDim.x)
            if (X[i] != X[i - 1] && fabs(X[i] - X[i - 1]) < DBL_EPSILON)
                atomicOr(&d.fs[f * kFsCount + kFsErr], kErrHullFuzzyOrder);
        // Stacks of sorted positions: lower in [0, n), upper in [n, 2n) of
        // the second buffer's doubles viewed as int32 (2n entries fit).
        int32_t* SL = reinterpret_cast<int32_t*>(d.hcx + BN + ob + base);
        int32_t* SU = SL + n;
        if (tid == 0) {
            int k = 0;
            for (int i = 0; i < n; ++i) {
                while (k > 1 && ref_ccw_or_collinear(X[SL[k - 2]], Y[SL[k - 2]], X[SL[k - 1]], Y[SL[k - 1]], X[i], Y[i]))
                    --k;
                SL[k++] = i;
            }
            nlo = k;
        } else if (tid == 1) {
            // The reference's upper loop continues on the lower stack with
            // t = k + 1; it never pops below the last point, so it is the
            // same chain started from that point.
            int k = 0;
            SU[k++] = n - 1;
            for (int i = n - 2; i >= 0; --i) {
                while (k > 1 && ref_ccw_or_collinear(X[SU[k - 2]], Y[SU[k - 2]], X[SU[k - 1]], Y[SU[k - 1]], X[i], Y[i]))
                    --k;
                SU[k++] = i;
            }
            nup = k;
        }
        __syncthreads();
        const int kl = nlo;
        const int kk = kl + nup - 2;  // indices.resize(k - 1)
        if (kk > n) {
            if (tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrHullCapacity);
        } else {
            for (int i = tid; i < kk; i += blockDim.x) d.hull_idx[ob + base + i] = I[i < kl ? SL[i] : SU[i - kl + 1]];
            if (tid == 0) {
                d.hull_count[ob + c] = static_cast<uint32_t>(kk);
                atomicAdd(&d.fs[f * kFsCount + kFsHullRounded], 1u);
            }
        }
        __syncthreads();
    }
}

// Hull vertices as cluster-local indices: rank of the obstacle index among
// the cluster's members (the gather order of processor.cpp:640-658).  Among
// members with exactly a vertex's coordinates the smallest index is reported:
// which of several identical points a chain keeps depends on how the rounds
// split a cluster (and the reference's own choice follows its unstable
// std::sort), so the coordinates are the defined result and the index is made
// canonical.  Coordinates come from the obstacle points, or for standalone
// lists from the copy staged in `list_xy` (list offset + list-local index).
// For the obstacle clusters the vertex coordinates (PointXY doubles, in hull
// order) also go to the second candidate buffer, free once the hulls are
// done, where the box stage reads its polygons.
//
// One workgroup per cluster: the hull (up to kLocalMax vertices) is staged in
// LDS with a hash of its coordinates; one pass over the members lowers each
// vertex's index to the smallest member index with its coordinates, a second
// pass bins every member by the number of (sorted) vertex indices not above
// it, and the prefix sums of the bins are the ranks.  Each thread takes four
// members at a time so their loads and searches overlap.  Larger hulls take
// one wavefront per vertex.
// Two launches split the clusters: those with at most kLocalSmallV hull
// vertices and kLocalSmallSize members (most of them) take one wavefront
// each, with 2.6 KiB of LDS, so a CU holds as many clusters as it holds
// wavefronts; the others take a 1024-thread workgroup with room for
// kLocalMax vertices, taken from a per-frame list the first launch writes.
constexpr int kLocalMax = 512;
constexpr int kLocalSmallV = 64;
constexpr int kLocalSmallSize = 1024;
constexpr int kLocalUnroll = 4;

// Hash of a coordinate pair that agrees with ==: +0.0 and -0.0 hash alike.
__device__ __forceinline__ uint32_t xy_hash(double x, double y) {
    const uint64_t a = static_cast<uint64_t>(__double_as_longlong(x + 0.0));
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(y + 0.0));
    uint64_t h = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
    h ^= h >> 29;
    return static_cast<uint32_t>(h);
}

template <int kLocalThreads, int kVMax, bool kSmall>
__global__ __launch_bounds__(kLocalThreads) void k_hull_local(DevBuf d, Layout L, const double2* list_xy) {
    constexpr int kLocalHash = 2 * kVMax;
    __shared__ int32_t canon[kVMax];
    __shared__ int32_t sorted[kVMax];
    __shared__ int16_t spos[kVMax];
    __shared__ int32_t table[kLocalHash];
    __shared__ double vx[kVMax];
    __shared__ double vy[kVMax];
    __shared__ uint32_t bins[kVMax + 1];
    int tile, f;
    xcd_tile_frame(tile, f);
    const int K = static_cast<int>(d.fs[f * kFsCount + kFsClusters]);
    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const int lane = lane_id();
    const size_t ob = (size_t)f * L.N;
    const size_t BN = (size_t)L.B * L.N;
    // The wavefront launch walks all clusters and lists the large ones; the
    // workgroup launch takes list entries tile, tile + gridDim.x, ...
    const int iters = kSmall ? K : static_cast<int>(d.fs[f * kFsCount + kFsHullLarge]);
    for (int j = tile; j < iters; j += gridDim.x) {
        const int c = kSmall ? j : d.hull_large[ob + j];
        const int V = static_cast<int>(d.hull_count[ob + c]);
        const uint32_t hoff = d.hull_off[ob + c];
        const uint32_t base = d.cl_off[ob + c];
        const int size = static_cast<int>(d.cl_size[ob + c]);
        const int32_t* members = d.cl_members + ob + base;
        if (kSmall && !(V <= kLocalSmallV && size <= kLocalSmallSize)) {
            if (tid == 0) d.hull_large[ob + atomicAdd(&d.fs[f * kFsCount + kFsHullLarge], 1u)] = c;
            continue;
        }
        // A cluster left without a hull ran out of the chunk rounds' and the
        // final workgroup's capacity.
        if (V == 0 && tid == 0) atomicOr(&d.fs[f * kFsCount + kFsErr], kErrHullCapacity);
        if (!kSmall && V > kVMax) {
            for (int i = w; i < V; i += kLocalThreads / kWave) {
                const int32_t idx = d.hull_idx[ob + base + i];
                double px, py;
                member_xy(d, list_xy, ob, base, idx, px, py);
                int32_t cn = idx;
                for (int q = lane; q < size; q += kWave) {
                    const int32_t m = members[q];
                    if (m >= cn) continue;
                    double mx, my;
                    member_xy(d, list_xy, ob, base, m, mx, my);
                    if (mx == px && my == py) cn = m;
                }
#pragma unroll
                for (int o = kWave / 2; o > 0; o >>= 1) cn = min(cn, __shfl_xor(cn, o, kWave));
                uint32_t cnt = 0;
                for (int q = lane; q < size; q += kWave) cnt += members[q] < cn ? 1u : 0u;
                cnt = wave_sum(cnt);
                if (lane == 0) {
                    d.hull_idx[ob + base + i] = cn;
                    d.hull_local[ob + hoff + i] = static_cast<int32_t>(cnt);
                    if (!list_xy) {
                        d.hcx[BN + ob + hoff + i] = px;
                        d.hcy[BN + ob + hoff + i] = py;
                    }
                }
            }
            continue;
        }
        const uint32_t hmask = static_cast<uint32_t>(kLocalHash - 1);
        for (int i = tid; i < kLocalHash; i += kLocalThreads) table[i] = -1;
        for (int i = tid; i <= V; i += kLocalThreads) bins[i] = 0;
        __syncthreads();
        for (int i = tid; i < V; i += kLocalThreads) {
            const int32_t idx = d.hull_idx[ob + base + i];
            double x, y;
            member_xy(d, list_xy, ob, base, idx, x, y);
            canon[i] = idx;
            vx[i] = x;
            vy[i] = y;
            if (!list_xy) {
                d.hcx[BN + ob + hoff + i] = x;
                d.hcy[BN + ob + hoff + i] = y;
            }
            uint32_t h = xy_hash(x, y) & hmask;  // load <= 1/2: a free slot exists
            while (atomicCAS(&table[h], -1, i) != -1) h = (h + 1) & hmask;
        }
        __syncthreads();
        // Smallest member index per vertex coordinates.
        for (int q0 = tid; q0 < size; q0 += kLocalThreads * kLocalUnroll) {
            int32_t m[kLocalUnroll];
            double mx[kLocalUnroll], my[kLocalUnroll];
#pragma unroll
            for (int u = 0; u < kLocalUnroll; ++u) {
                const int q = q0 + u * kLocalThreads;
                m[u] = q < size ? members[q] : -1;
            }
#pragma unroll
            for (int u = 0; u < kLocalUnroll; ++u)
                if (m[u] >= 0) member_xy(d, list_xy, ob, base, m[u], mx[u], my[u]);
#pragma unroll
            for (int u = 0; u < kLocalUnroll; ++u) {
                if (m[u] < 0) continue;
                for (uint32_t h = xy_hash(mx[u], my[u]) & hmask;; h = (h + 1) & hmask) {
                    const int e = table[h];
                    if (e < 0) break;
                    if (vx[e] == mx[u] && vy[e] == my[u] && m[u] < canon[e]) atomicMin(&canon[e], m[u]);
                }
            }
        }
        __syncthreads();
        // Sorted order of the vertex indices (ties, i.e. one point reported
        // twice, in vertex order).
        for (int i = tid; i < V; i += kLocalThreads) {
            const int32_t ci = canon[i];
            int pos = 0;
            for (int j = 0; j < V; ++j) {
                const int32_t cj = canon[j];
                pos += (cj < ci || (cj == ci && j < i)) ? 1 : 0;
            }
            sorted[pos] = ci;
            spos[i] = static_cast<int16_t>(pos);
            d.hull_idx[ob + base + i] = ci;
        }
        __syncthreads();
        // Member m lands in bin b = number of sorted indices <= m; it is below
        // sorted[k] exactly when b <= k.  Neighbouring lanes mostly share a
        // bin: each run of equal bins adds its length once.
        int steps = 0;
        while ((1 << steps) <= V) ++steps;
        // Wave-uniform trip count: the run detection below reads all lanes.
        for (int q0 = w * kWave + lane; q0 - lane < size; q0 += kLocalThreads * kLocalUnroll) {
            int32_t m[kLocalUnroll];
            int lo[kLocalUnroll], hi[kLocalUnroll];
#pragma unroll
            for (int u = 0; u < kLocalUnroll; ++u) {
                const int q = q0 + u * kLocalThreads;
                m[u] = q < size ? members[q] : -1;
                lo[u] = 0;
                hi[u] = V;
            }
            for (int st = 0; st < steps; ++st) {
#pragma unroll
                for (int u = 0; u < kLocalUnroll; ++u) {
                    if (lo[u] < hi[u]) {
                        const int mid = (lo[u] + hi[u]) >> 1;
                        if (sorted[mid] <= m[u])
                            lo[u] = mid + 1;
                        else
                            hi[u] = mid;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kLocalUnroll; ++u) {
                const int b = m[u] >= 0 ? lo[u] : -1;
                const int prev = __shfl_up(b, 1, kWave);
                const bool start = lane == 0 || b != prev;
                const uint64_t starts = __ballot(start);
                if (start && b >= 0) {
                    const uint64_t above = starts & ~((2ull << lane) - 1ull);
                    const int next = above ? __ffsll(static_cast<unsigned long long>(above)) - 1 : kWave;
                    atomicAdd(&bins[b], static_cast<uint32_t>(next - lane));
                }
            }
        }
        __syncthreads();
        if (w == 0) {
            uint32_t carry = 0;
            for (int i0 = 0; i0 < V; i0 += kWave) {
                const int i = i0 + lane;
                uint32_t v = i < V ? bins[i] : 0u;
#pragma unroll
                for (int o = 1; o < kWave; o <<= 1) {
                    const uint32_t t = __shfl_up(v, o, kWave);
                    if (lane >= o) v += t;
                }
                if (i < V) bins[i] = carry + v;  // members below sorted[i]
                carry += __shfl(v, kWave - 1, kWave);
            }
        }
        __syncthreads();
        for (int i = tid; i < V; i += kLocalThreads) d.hull_local[ob + hoff + i] = static_cast<int32_t>(bins[spos[i]]);
        __syncthreads();
    }
}

// Standalone lists: per point, its list and list-local index; per list, the
// offset and size (cl_off / cl_size) and identity members.
__global__ void k_hull_lists_init(DevBuf d, Layout L, const int32_t* offsets, int K, int T) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        d.fs[kFsClusters] = static_cast<uint32_t>(K);
        d.fs[kFsErr] = 0;
        d.fs[kFsHullInexact] = 0;
        d.fs[kFsHullRounded] = 0;
        d.fs[kFsHullLarge] = 0;
    }
    if (t < K) {
        d.cl_off[t] = static_cast<uint32_t>(offsets[t]);
        d.cl_size[t] = static_cast<uint32_t>(offsets[t + 1] - offsets[t]);
        d.hc_count[t] = d.cl_size[t];  // round-0 input: the whole list
    }
    if (t < T) {
        int lo = 0, hi = K;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (offsets[mid] <= t)
                lo = mid + 1;
            else
                hi = mid;
        }
        const int local = t - offsets[lo - 1];
        d.hci[t] = local;
        d.cl_members[t] = local;
    }
}

void launch_hull_lists(lp_context* c, int num_lists, int total_points) {
    const int n = std::max(num_lists, total_points);
    const int32_t* offs = reinterpret_cast<const int32_t*>(c->d.hc_choff);  // staged by the caller
    hipLaunchKernelGGL(k_hull_lists_init, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d, c->L, offs,
                       num_lists, total_points);
    launch_hull(c, 1, reinterpret_cast<const double2*>(c->d.dsort));
}

void launch_hull(lp_context* c, int B, const double2* list_xy) {
    DevBuf& d = c->d;
    const Layout& L = c->L;
    hipStream_t st = c->stream;
    // Stage marks around the single-kernel steps ("hull_round" three
    // times), so profiled runs time each round launch on its own.
    // Chunk wavefronts loop over the frame's chunks: about 2048 workgroups
    // in all, at most 512 per frame; workgroups beyond that would mostly
    // find no chunk and only cost dispatch.
    const int round_grid = std::min(512, std::max(8, 2048 / B));
    for (int r = 0; r < kHullRounds; ++r) {
        stage_mark(c, "hull");
        hipLaunchKernelGGL(k_hull_plan, dim3(B), dim3(kPlanThreads), 0, st, d, L, r);
        stage_mark(c, "hull_round");
        hipLaunchKernelGGL(k_hull_round, dim3(round_grid, B), dim3(256), 0, st, d, L, r);
    }
    stage_mark(c, "hull_big");
    hipLaunchKernelGGL(k_hull_big, dim3(16, B), dim3(1024), 0, st, d, L, kHullRounds);
    stage_mark(c, "hull_ref");
    hipLaunchKernelGGL(k_hull_ref, dim3(16, B), dim3(256), 0, st, d, L, list_xy);
    stage_mark(c, "hull");
    launch_scan_small(c, B, d.hull_count, d.hull_off, L.N, kCountSlot, kFsClusters, kFsHullTotal);
    stage_mark(c, "hull_local");
    // About 16384 wavefronts for the small clusters (about one cluster each
    // on 128 frames), at least 8 per frame.
    const int small_grid = std::min(1024, std::max(8, 16384 / B));
    hipLaunchKernelGGL((k_hull_local<kWave, kLocalSmallV, true>), dim3(small_grid, B), dim3(kWave), 0, st, d, L,
                       list_xy);
    // The large clusters, listed by the launch above, are a few dozen per
    // frame at most: 16 to 256 workgroups per frame, so each takes about one.
    const int large_grid = std::min(256, std::max(16, 2048 / B));
    hipLaunchKernelGGL((k_hull_local<1024, kLocalMax, false>), dim3(large_grid, B), dim3(1024), 0, st, d, L,
                       list_xy);
}

}  // namespace lp
