// Run walk of the pair kernel: the pixel pairs of one pixel i of delta1 with the pixels j of
// a monotone delta2, taken as runs of consecutive j that share one bin.
//
// Along a forest with strictly increasing r_comov and dist_m, the reference's r_par bin
// (py/picca/cf.py:355-376) of the pair (i, j) falls as j approaches r_comov1[i] and rises
// after it, and its r_trans bin never falls: every FP64 operation on the way is monotone.
// So the pairs of pixel i split into runs of consecutive j with one (side, r_par bin,
// r_trans bin) key, and a run's contribution to compute_xi's five sums and the pair count
// is the difference of two entries of prefix sums over delta2.  The walk finds each run's
// end with a few probes (a guess from the local pixel spacing, then steps and bisection),
// so its cost grows with the number of runs, not of pairs.
//
// Exactness: a probe takes the bin coordinates with a reciprocal multiply and redoes the
// pixel with the reference's arithmetic when a coordinate lies within the rounding margin
// of an integer, so every probed key is the reference's.  Keys are monotone, so the pixels
// between two probes with one key share it.  A pixel within the margin of a bin edge or
// range limit is the first or last pixel of its run whenever consecutive pixels of delta2
// are more than four margins apart in both coordinates (the kernel walks only such pairs),
// and those pixels get the reference's edge, overflow and wrap-around treatment one by one.
//
// The header is compiled into the kernel and into a CPU harness of the tests, which runs
// the same lane program under bounds checks; `Mem` supplies the memory accesses.
#pragma once

#include <math.h>
#include <stdint.h>

#include "engine.h"

#if defined(__HIPCC__)
#define PICCA_HD __host__ __device__
#else
#define PICCA_HD
#endif

namespace picca {

// Entry k of a forest's prefix sums covers its pixels [0, k): pixels with w != 0, and the
// sums of w, w * delta, w * (r_comov - r_c), w * (dist_m - d_c) and w * z, where r_c and
// d_c are r_comov and dist_m of the forest's pixel L / 2.  A forest of L pixels has L + 1
// entries, starting at entry offsets[b] + b.
struct alignas(16) PrefixRecord {
    double c, w, x, r, d, z;
};
static_assert(sizeof(PrefixRecord) == 48, "PrefixRecord layout");

// Host side: the prefix entries of one forest of n pixels (n + 1 entries) and the smallest
// step of r_comov and of dist_m between consecutive pixels (-1 unless both increase
// strictly; a forest of fewer than two pixels has no step and gets +1e300).  dw = delta * w,
// the same products the pair loops use.
inline void build_walk_forest(const double *r, const double *d, const double *z,
                              const double *w, const double *dw, int64_t n,
                              PrefixRecord *pre, double &spc_r, double &spc_d) {
    const double rc = n > 0 ? r[n / 2] : 0.0;
    const double dc = n > 0 ? d[n / 2] : 0.0;
    PrefixRecord acc = {0, 0, 0, 0, 0, 0};
    pre[0] = acc;
    spc_r = 1e300;
    spc_d = 1e300;
    for (int64_t k = 0; k < n; ++k) {
        if (w[k] != 0) acc.c += 1.0;
        acc.w += w[k];
        acc.x += dw[k];
        acc.r += w[k] * (r[k] - rc);
        acc.d += w[k] * (d[k] - dc);
        acc.z += w[k] * z[k];
        pre[k + 1] = acc;
        if (k > 0) {
            const double sr = r[k] - r[k - 1];
            const double sd = d[k] - d[k - 1];
            spc_r = sr < spc_r ? sr : spc_r;
            spc_d = sd < spc_d ? sd : spc_d;
        }
    }
    if (!(spc_r > 0) || !(spc_d > 0)) {
        spc_r = -1.0;
        spc_d = -1.0;
    }
}

// Per (forest pair, chunk of delta1): wave-uniform constants of the walk.
struct WalkTask {
    double cos_half, sin_half;
    double cp, st;       // cos(ang/2) * np / span_p and sin(ang/2) * nt / r_trans_max
    double rc, dc;       // r_c and d_c of delta2's prefix sums
    double q0;           // first estimate of delta2 pixels per unit of the r_par coordinate
    double r_lo;         // r_comov of delta2's pixel j_lo
    int32_t j_lo, j_hi;  // pruned pixel window of delta2
};

// Pixel i of delta1 (weight != 0).
struct WalkLane {
    double r1, d1, zh1, w1, dw1;  // zh1 = z1 / 2, dw1 = delta1 * w1
    double phase;                 // in [0, 1): where in its r_par range the walk starts
};

// Key of a pixel pair: m is the signed r_par run index, m = bin for r_comov2 >= r_comov1 and
// -bin - 1 below it, clamped to [-np - 1, np] (|m| beyond np - 1 is rejected); c is the
// r_trans bin clamped to nt (nt is rejected, or a wrap-around pair).
struct WalkKey {
    int32_t m, c;
};

// Everything the reference does with one pixel pair (standard binning, auto-correlation,
// r_par_min == 0, no cuts), py/picca/cf.py:355-376, with its key.
struct WalkExact {
    int32_t m, c;
    bool edge;      // within edge_tol of a bin edge or of a range limit
    bool over;      // accepted, flat bin index >= np * nt (picca writes out of bounds)
    bool wrap;      // accepted with its r_trans bin rounded to nt: the flat index is
                    // column 0 of the next r_par row, where the pair is added
    int32_t pbin;   // padded histogram bin of a wrap pair
    double r_par, r_trans;
};

PICCA_HD inline double walk_edge_distance(double t, double width) {
    const double frac = t - floor(t);
    const double lo = frac * width;
    const double hi = (1.0 - frac) * width;
    return lo < hi ? lo : hi;
}

PICCA_HD inline WalkExact walk_exact(const XiConst &C, double r1, double d1, double r2,
                                     double d2, double cos_half, double sin_half) {
    WalkExact e;
    const double rps = (r1 - r2) * cos_half;
    const double r_par = fabs(rps);
    const double r_trans = (d1 + d2) * sin_half;
    const double t_par = r_par / C.span_p * C.np_d;
    const double t_trans = r_trans / C.rtmax * C.nt_d;
    const double f_par = floor(t_par);
    const double f_trans = floor(t_trans);
    const double row = fmin(f_par, C.np_d);
    e.m = (int32_t)(rps > 0 ? -row - 1.0 : row);
    e.c = (int32_t)fmin(f_trans, C.nt_d);
    const bool out_p = r_par >= C.rpmax;
    const bool out_t = r_trans >= C.rtmax;
    const bool out = out_p | out_t;
    const bool edge_out = (out_p & (r_par - C.rpmax < C.edge_tol)) |
                          (out_t & (r_trans - C.rtmax < C.edge_tol));
    const bool edge_in = (r_par != 0 && walk_edge_distance(t_par, C.width_p) < C.edge_tol) ||
                         (r_trans != 0 && walk_edge_distance(t_trans, C.width_t) < C.edge_tol);
    e.edge = out ? edge_out : edge_in;
    const double bin_f = f_trans + C.nt_d * f_par;
    e.over = !out & !(bin_f < C.nb_d);
    e.wrap = !out & !e.over & (f_trans >= C.nt_d);
    e.pbin = e.wrap ? (int32_t)(C.ntp_d * (f_par + 1.0)) : 0;
    e.r_par = r_par;
    e.r_trans = r_trans;
    return e;
}

// Key from the reciprocal-multiply coordinates; `near` flags a coordinate within the
// rounding margin of an integer, where only walk_exact gives the reference's key.
PICCA_HD inline WalkKey walk_fast_key(const XiConst &C, const WalkTask &T, double r1, double d1,
                                      double r2, double d2, double &t_par, bool &near) {
    t_par = (r2 - r1) * T.cp;
    const double t_tr = (d1 + d2) * T.st;
    const double fp = floor(t_par);
    const double ft = floor(t_tr);
    near = fabs(t_par - fp - 0.5) > C.half_p || fabs(t_tr - ft - 0.5) > C.half_t;
    WalkKey k;
    k.m = (int32_t)fmin(fmax(fp, -C.np_d - 1.0), C.np_d);
    k.c = (int32_t)fmin(ft, C.nt_d);
    return k;
}

template <class Mem>
PICCA_HD inline WalkKey walk_key(const XiConst &C, const WalkTask &T, const WalkLane &L,
                                 Mem &mem, int p, double &t_par, bool &near) {
    double r2, d2;
    mem.rd(p, r2, d2);
    WalkKey k = walk_fast_key(C, T, L.r1, L.d1, r2, d2, t_par, near);
    if (near) {
        const WalkExact e = walk_exact(C, L.r1, L.d1, r2, d2, T.cos_half, T.sin_half);
        k.m = e.m;
        k.c = e.c;
    }
    return k;
}

// The reference's treatment of a pixel within the margin of an edge: edge and overflow
// counts, and the wrap-around pair, which no run holds.  Pixels with w == 0 (either sign)
// are skipped before any arithmetic, as in the reference.
template <class Mem>
PICCA_HD inline void walk_special(const XiConst &C, const WalkTask &T, const WalkLane &L,
                                  Mem &mem, int p) {
    const double w2 = mem.w(p);
    if (w2 == 0) return;
    double r2, d2;
    mem.rd(p, r2, d2);
    const WalkExact e = walk_exact(C, L.r1, L.d1, r2, d2, T.cos_half, T.sin_half);
    if (e.edge) mem.count_edge();
    if (e.over) mem.count_overflow();
    if (e.wrap) {
        const double w12 = L.w1 * w2;
        const double z = L.zh1 + mem.z(p) * 0.5;
        mem.add(e.pbin, w12, L.dw1 * mem.dw(p), e.r_par * w12, e.r_trans * w12, z * w12, 1u);
    }
}

// Adds the run [j, jn) of key k: prefix differences times the lane's pixel.
template <class Mem>
PICCA_HD inline void walk_flush(const XiConst &C, const WalkTask &T, const WalkLane &L, Mem &mem,
                                WalkKey k, const PrefixRecord &ps, const PrefixRecord &pe) {
    const double n = pe.c - ps.c;
    if (!(n > 0)) return;
    const int row = k.m >= 0 ? k.m : -k.m - 1;
    const double dw = pe.w - ps.w;
    const double dx = pe.x - ps.x;
    const double dr = pe.r - ps.r;
    const double dd = pe.d - ps.d;
    const double dz = pe.z - ps.z;
    // sum over the run of w2 (r2 - r1) and w2 (d1 + d2), from the centred sums
    const double s_r = dr - (L.r1 - T.rc) * dw;
    const double s_d = (L.d1 + T.dc) * dw + dd;
    const double rp = T.cos_half * (L.w1 * (k.m >= 0 ? s_r : -s_r));
    const double rt = T.sin_half * (L.w1 * s_d);
    const double z = L.w1 * (L.zh1 * dw + 0.5 * dz);
    mem.add(row * C.ntp + k.c, L.w1 * dw, L.dw1 * dx, rp, rt, z, (uint32_t)n);
}

// First guess of the end of the run that starts at pixel j with key k: the pixel where the
// r_par coordinate reaches the run's upper edge at q pixels per unit, in (j, j_end].
PICCA_HD inline int walk_guess(WalkKey k, double t_j, double q, int j, int j_end) {
    double ahead = ceil(((double)(k.m + 1) - t_j) * q);
    ahead = fmin(fmax(ahead, 1.0), (double)(j_end - j));
    return j + (int)ahead;
}

PICCA_HD inline bool walk_same(WalkKey a, WalkKey b) { return a.m == b.m && a.c == b.c; }

// r_comov2 beyond r_par_max on the far side, or r_trans beyond r_trans_max: every later
// pixel is rejected too
PICCA_HD inline bool walk_terminal(const XiConst &C, WalkKey k) {
    return k.m >= C.np || k.c >= C.nt;
}

// All pairs of lane pixel L with delta2's window [T.j_lo, T.j_hi).
//
// The window is walked as [s, j_hi) then [j_lo, s).  Every lane meets the bins in the same
// order along j and the lanes of a wave move at about the same pace, so walks that all
// began at j_lo would add to a few neighbouring bins at a time; the split point s sits at
// the fraction `phase` of the lane's r_par range [r1 - r_par_max / cos, r1 + r_par_max /
// cos] (placed with the window's pixel density; any s is exact), so the lanes of a wave
// add to different bins.
//
// One loop, one memory round trip per pass: every pass loads pixels pa, pb (their r_comov
// and dist_m) and the prefix entry at pb, evaluates both keys, and then updates the lane's
// state with short selects, so the lanes of a wave stay on one instruction stream whatever
// state each is in:
//   * start of a segment at j (pa = pb = j): the run's key, its prefix entry, the first
//     guess of its end;
//   * test of a candidate end g of the run [j, ...) (pa = g - 1, pb = g): pixel g - 1 with
//     the run's key and pixel g without it close the run at g; otherwise the bracket
//     (lo, hi) of the end shrinks (a step back after the first overshoot, then bisection;
//     galloping forward after an undershoot).
// A run whose end is guessed right closes in one pass.
template <class Mem>
PICCA_HD inline void walk_lane(const XiConst &C, const WalkTask &T, const WalkLane &L, Mem &mem) {
    if (T.j_lo >= T.j_hi) return;
    const double r_start = L.r1 + (2.0 * L.phase - 1.0) * (C.np_d / T.cp);
    double s = (double)T.j_lo + (r_start - T.r_lo) * T.cp * T.q0;
    s = fmin(fmax(s, (double)T.j_lo), (double)T.j_hi);
    const int split = (int)s;
    // first segment [split, j_hi); an empty one goes straight to [j_lo, split)
    bool second = split >= T.j_hi;
    int seg_end = second ? split : T.j_hi;
    int j = second ? T.j_lo : split;
    bool active = j < seg_end;
    bool starting = true;
    WalkKey k = {0, 0};
    double t_j = 0.0, q = T.q0;
    bool near_j = false, near_lo = false;
    PrefixRecord ps = {0, 0, 0, 0, 0, 0};
    int lo = j, hi = seg_end, g = j, misses = 0, step = 1;
    const int last = T.j_hi > 0 ? T.j_hi - 1 : 0;   // pixel reads stay inside the window
    while (active) {
        const int pa = starting ? j : g - 1;
        const int pb = starting ? j : g;
        double ta, tb;
        bool na, nb;
        const WalkKey ka = walk_key(C, T, L, mem, pa, ta, na);
        const WalkKey kb = walk_key(C, T, L, mem, pb < last ? pb : last, tb, nb);
        const PrefixRecord pg = mem.prefix(pb);
        mem.count_probe();
        bool open = false;     // a run starts at pb
        bool end_seg = false;  // the current segment is done
        if (starting) {
            open = true;
        } else {
            const bool has1 = g - 1 > lo;
            const bool has2 = g < seg_end;
            const bool same1 = !has1 || walk_same(ka, k);
            const bool same2 = has2 && walk_same(kb, k);
            if (!same1) {                  // the run ends at g - 1 or before
                hi = g - 1;
                g = misses == 0 ? hi : lo + 1 + ((hi - lo) >> 1);
                ++misses;
            } else if (same2) {            // the run goes on past g
                lo = g;
                near_lo = nb;
                g = g + step < hi ? g + step : hi;
                step <<= 1;
            } else {                       // the run is [j, g)
                if ((has1 ? na : near_lo) && g - 1 > j) walk_special(C, T, L, mem, g - 1);
                mem.count_run(pg.c - ps.c);
                if (k.m >= -C.np) walk_flush(C, T, L, mem, k, ps, pg);
                if (has2) {
                    if (g - j >= 3 && tb > t_j) q = (double)(g - j) / (tb - t_j);
                    open = true;
                } else {
                    end_seg = true;
                }
            }
        }
        if (open) {                        // a run starts at pixel pb with key kb
            j = pb;
            k = kb;
            t_j = tb;
            near_j = nb;
            ps = pg;
            starting = false;
            if (near_j) walk_special(C, T, L, mem, j);
            if (walk_terminal(C, k)) {
                end_seg = true;
            } else {
                lo = j;
                hi = seg_end;
                near_lo = near_j;
                misses = 0;
                step = 1;
                g = walk_guess(k, t_j, q, j, seg_end);
            }
        }
        if (end_seg) {
            if (!second && T.j_lo < split) {
                second = true;
                seg_end = split;
                j = T.j_lo;
                starting = true;
            } else {
                active = false;
            }
        }
    }
}

}  // namespace picca
