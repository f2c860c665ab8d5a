// Run walk of the pair kernel: the pixel pairs of one pixel i of delta1 with the pixels j of
// a monotone delta2, taken as runs of consecutive j that share one bin.
//
// Along a forest with strictly increasing r_comov and dist_m, the reference's r_par bin
// (py/picca/cf.py:355-376) of the pair (i, j) falls as j approaches r_comov1[i] and rises
// after it, and its r_trans bin never falls: every FP64 operation on the way is monotone.
// So the pairs of pixel i split into runs of consecutive j with one (side, r_par bin,
// r_trans bin) key, and a run's contribution to compute_xi's five sums and the pair count
// is the difference of two entries of prefix sums over delta2.  The walk finds each run's
// end with a few probes (a guess from the window's pixel density, then steps and
// bisection), so its cost grows with the number of runs, not of pairs.
//
// Exactness: a probe takes the bin coordinates with reciprocal multiplies and redoes the
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
    double q0;           // mean delta2 pixels per unit of the r_par coordinate in the window
    int32_t j_lo, j_hi;  // pruned pixel window of delta2
};

// Pixel i of delta1 (weight != 0).
struct WalkLane {
    double r1, d1, zh1, w1, dw1;  // zh1 = z1 / 2, dw1 = delta1 * w1
};

// Everything the reference does with one pixel pair (standard binning, auto-correlation,
// r_par_min == 0, no cuts), py/picca/cf.py:355-376, with its key: m is the signed r_par run
// index, m = bin for r_comov2 >= r_comov1 and -bin - 1 below it, clamped to [-np - 1, np]
// (|m| beyond np - 1 is rejected); c is the r_trans bin clamped to nt (nt is rejected, or a
// wrap-around pair).
struct WalkExact {
    int32_t m, c;
    bool edge;      // within edge_tol of a bin edge, or outside the range but within
                    // edge_tol of it in both coordinates
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
    // outside the range, but within edge_tol of it in both coordinates
    const bool edge_out = (r_par - C.rpmax < C.edge_tol) & (r_trans - C.rtmax < C.edge_tol);
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

// Histogram addition of the run [j, jn) of key (m, c): prefix differences times the lane's
// pixel.  False for an empty run.  No branch, and finite for any finite state.
struct RunSums {
    int bin;
    double we, xi, rp, rt, z;
    uint32_t n;
};

PICCA_HD inline bool walk_run_sums(const XiConst &C, const WalkTask &T, const WalkLane &L, int m,
                                   int c, const PrefixRecord &ps, const PrefixRecord &pe,
                                   RunSums &o) {
    const double n = pe.c - ps.c;  // pixels of the run with w2 != 0 (>= 0)
    const double dw = pe.w - ps.w;
    const double dx = pe.x - ps.x;
    const double dr = pe.r - ps.r;
    const double dd = pe.d - ps.d;
    const double dz = pe.z - ps.z;
    // sum over the run of w2 (r2 - r1) and w2 (d1 + d2), from the centred sums
    const double s_r = dr - (L.r1 - T.rc) * dw;
    const double s_d = (L.d1 + T.dc) * dw + dd;
    o.bin = (m >= 0 ? m : -m - 1) * C.ntp + c;
    o.we = L.w1 * dw;
    o.xi = L.dw1 * dx;
    o.rp = T.cos_half * (L.w1 * (m >= 0 ? s_r : -s_r));
    o.rt = T.sin_half * (L.w1 * s_d);
    o.z = L.w1 * (L.zh1 * dw + 0.5 * dz);
    o.n = (uint32_t)fmax(n, 0.0);
    return n > 0;
}

// Reciprocal of x > 0 within 2.5e-4 relative, the same on the host and the device: the
// seed is the bit pattern of 2.0^1023 (0x7FE0...) minus that of x, within 12.5 % of 1 / x
// for a normal x, then two Newton steps y (2 - x y).  For the guess of a run's end only.
PICCA_HD inline double walk_rcp_guess(double x) {
    const uint64_t b = 0x7FE0000000000000ull - __builtin_bit_cast(uint64_t, x);
    double y = __builtin_bit_cast(double, b);
    y = y * fma(-x, y, 2.0);
    y = y * fma(-x, y, 2.0);
    return y;
}

// Key of pixel p from its coordinates relative to the upper edges of the run key (m, c):
// a = (r2 - r1) cp - (m + 1) and b = (d1 + d2) st - (c + 1), so the pixel's r_par run index
// is m + 1 + floor(a) and its r_trans bin c + 1 + floor(b), clamped as in walk_exact.
PICCA_HD inline void walk_key_from(const XiConst &C, int m, int c, double a, double b,
                                   int &mp, int &cp) {
    // |a| and |b| stay within a few times the bin counts inside the pruned window
    const int fm = (int)floor(a) + (m + 1);
    const int fc = (int)floor(b) + (c + 1);
    mp = fm < -C.np - 1 ? -C.np - 1 : (fm > C.np ? C.np : fm);
    cp = fc > C.nt ? C.nt : fc;
}

// Whether a task takes the run walk: delta2 increases strictly with consecutive pixels more
// than four rounding margins apart in both bin coordinates (spc_r, spc_d: its smallest steps
// of r_comov and dist_m), and, under the automatic schedule, the window holds at least
// XI_WALK_MIN_DENSITY pixels per unit of the r_par coordinate (q0).
PICCA_HD inline bool walk_task(const XiConst &C, double spc_r, double spc_d, double cp, double st,
                               double q0) {
    return (C.flags & XI_WALK) && spc_r > 0 && spc_d > 0 && cp * spc_r > 4.0 * (0.5 - C.half_p) &&
           st * spc_d > 4.0 * (0.5 - C.half_t) &&
           (!(C.flags & XI_WALK_AUTO) || q0 >= XI_WALK_MIN_DENSITY);
}

// r_comov2 beyond r_par_max on the far side, or r_trans beyond r_trans_max: every later
// pixel is rejected too
PICCA_HD inline bool walk_terminal(const XiConst &C, int m, int c) {
    return m >= C.np || c >= C.nt;
}

// All pairs of lane pixel L with delta2's window [T.j_lo, T.j_hi).
//
// Every lane starts at j_lo and walks its runs in order, so the lanes of a wave (consecutive
// pixels of delta1) probe neighbouring pixels of delta2 at the same time and their loads
// share cache lines.  (Starting each lane at a different point of its r_par range, so that
// the lanes add to different bins, and prefetching two runs ahead were both measured
// slower.)
//
// One loop, one memory round trip per pass: every pass tests pixels pa, pb (their r_comov
// and dist_m) and the prefix entry at pb against the thresholds of the current run key
// (Ka = r1 cp + m + 1, Kb = c + 1 - d1 st: a pixel at or after the run's first pixel keeps
// the key iff r2 cp < Ka and d2 st < Kb):
//   * start of the walk at j (pa = pb = j): the run's key, its prefix entry, the first
//     guess of its end;
//   * test of a candidate end g of the run [j, ...) (pa = g - 1, pb = g): pixel g - 1 with
//     the run's key and pixel g without it close the run at g, and pixel g opens the next
//     run with its own key; otherwise the bracket (lo, hi) of the end shrinks (a step back
//     after the first overshoot, then bisection; galloping forward after an undershoot).
// The next pass's loads are issued before the closed run is added to the histogram; the
// addition is made by all lanes of the wave together (mem.flush), so every lane stays in
// the loop until the wave's last lane is done (`alive` false: a lane without a pixel, or
// one past its last run).  The guess of a run's end is its first pixel plus the distance
// to the run's upper r_par edge times the pixel density where the previous run closed (the
// window's mean for the first run), so a run is usually closed in one pass.
// A coordinate within the rounding margin of an integer is redone with the reference's
// arithmetic (walk_exact); such a pixel is the first or last pixel of its run and gets the
// reference's edge, overflow and wrap-around treatment (walk_special).
template <class Mem>
PICCA_HD inline void walk_lane(const XiConst &C, const WalkTask &T, const WalkLane &L, Mem &mem,
                               bool alive) {
    const int j_hi = T.j_hi;
    if (T.j_lo >= j_hi) return;  // wave-uniform
    const int last = j_hi - 1;  // pixel reads stay inside the window
    const double r1cp = L.r1 * T.cp;
    const double d1st = L.d1 * T.st;
    // Before the first run opens the key is (-1, -1), which no pixel has (r_trans >= 0
    // here), so the thresholds give each pixel's own coordinates.
    int m = -1, c = -1;
    double Ka = r1cp, Kb = -d1st;
    int j = T.j_lo;
    bool starting = true;
    PrefixRecord ps = {0, 0, 0, 0, 0, 0};
    int lo = j, hi = j_hi, g = j, misses = 0, step = 1;
    bool near_lo = false;
    // the pass's probes: pixels pa and pb and the prefix entry at pb, loaded at the end of
    // the previous pass so that the loads overlap that pass's histogram additions
    int pa = j, pb = j;
    double ra, da, rb, db;
    mem.rd(pa, ra, da);
    mem.rd(pb, rb, db);
    PrefixRecord pg = mem.prefix(pb);
    for (;;) {
        // the sums of the run [j, g) of the current key, added by mem.flush with every lane
        // of the wave if the run closes in this pass (`on`); computed by every lane, so that
        // the values the flush moves between lanes are finite
        RunSums run;
        const bool nonempty = walk_run_sums(C, T, L, m, c, ps, pg, run);
        mem.pin(run);
        bool on = false;
        if (alive) {
            mem.count_probe();
            const double aA = fma(ra, T.cp, -Ka), bA = fma(da, T.st, -Kb);
            const double aB = fma(rb, T.cp, -Ka), bB = fma(db, T.st, -Kb);
            // pixel pb: within the margin of any bin edge; pixel pa: of the run's upper edges
            const bool nearB =
                (fabs(aB - rint(aB)) < C.fast_p) || (fabs(bB - rint(bB)) < C.fast_t);
            const bool nearA = (fabs(aA) < C.fast_p) || (fabs(bA) < C.fast_t);
            int mB, cB;
            walk_key_from(C, m, c, aB, bB, mB, cB);
            if (nearB) {
                const WalkExact e = walk_exact(C, L.r1, L.d1, rb, db, T.cos_half, T.sin_half);
                mB = e.m;
                cB = e.c;
            }
            bool inA = (aA < 0) && (bA < 0);
            if (nearA && !starting) {
                const WalkExact e = walk_exact(C, L.r1, L.d1, ra, da, T.cos_half, T.sin_half);
                inA = (e.m == m) && (e.c == c);
            }
            // outcome of the pass
            const bool has1 = g - 1 > lo;    // pa is not yet known to have the run's key
            const bool has2 = g < j_hi;      // pb lies inside the window
            const bool same1 = !has1 || inA;
            const bool same2 = has2 && (mB == m) && (cB == c);
            const bool miss = !starting && !same1;           // the run ends at g - 1 or before
            const bool gallop = !starting && same1 && same2; // the run goes on past g
            const bool close = !starting && same1 && !same2; // the run is [j, g)
            const bool open = starting || (close && has2);   // a run starts at pb
            const bool terminal = open && walk_terminal(C, mB, cB);
            const bool done = (close && !has2) || terminal;
            const bool special_last = close && (has1 ? nearA : near_lo) && g - 1 > j;
            const int p_last = g - 1;
            const bool special_first = open && nearB;
            const int p_first = pb;
            // the closing run [j, g): prefix entries ps and pg (the entry at g)
            if (close) {
                mem.count_run(run.n);
                on = nonempty && m >= -C.np;
            }
            // next state
            const int g_miss = misses == 0 ? g - 1 : lo + 1 + ((g - 1 - lo) >> 1);
            const int g_gallop = g + step < hi ? g + step : hi;
            if (miss) {
                hi = g - 1;
                g = g_miss;
                ++misses;
            }
            if (gallop) {
                lo = g;
                near_lo = nearB;
                g = g_gallop;
                step <<= 1;
            }
            if (open && !terminal) {
                // pixels per unit of the r_par coordinate: from the spacing of pixels g - 1
                // and g where a run closed here, else the window's mean
                const double q = starting ? T.q0 : walk_rcp_guess(T.cp * (rb - ra));
                j = pb;
                m = mB;
                c = cB;
                ps = pg;
                starting = false;
                Ka = r1cp + (double)(m + 1);
                Kb = (double)(c + 1) - d1st;
                lo = j;
                hi = j_hi;
                near_lo = nearB;
                misses = 0;
                step = 1;
                // pixels to the run's upper r_par edge
                const double ahead = ceil(-fma(rb, T.cp, -Ka) * q);
                g = j + (int)fmin(fmax(ahead, 1.0), (double)(j_hi - j));
            }
            if (!done) {
                pa = g - 1;
                pb = g;
            }
            if (special_last) walk_special(C, T, L, mem, p_last);
            if (special_first) walk_special(C, T, L, mem, p_first);
            if (done) alive = false;
        }
        // the next pass's probes, issued before the histogram additions (a lane that is done
        // reloads its last probes, which lie inside the window)
        const bool more = mem.any(alive);
        mem.rd(pa, ra, da);
        mem.rd(pb < last ? pb : last, rb, db);
        pg = mem.prefix(pb);
        mem.flush(on, run);
        if (!more) break;
    }
}

}  // namespace picca
