// Host-side neighbour candidates for fill_neighs (py/picca/cf.py:82-135).
//
// The reference asks healpy.query_disc(inclusive=True) for the pixels that touch a disc
// of radius ang_max, keeps the forests of those pixels, and then applies the exact test
// ang < ang_max with the angle of QSO.get_angle_between.  Any superset of the true
// neighbours gives the same result after that test, so this code replaces the disc
// query by a uniform 3-D grid over the unit vectors with cells as wide as the largest
// chord, and keeps every target whose dot product reaches cos_min (the caller passes
// cos(ang_max) minus a margin) or that lies within the 2-arcsec small-angle box where
// the reference switches to the flat-sky formula.  The dot product is evaluated as
// x2*x1 + y2*y1 + z2*z1 without contraction, bit-for-bit the value the reference
// feeds to arccos; the caller takes the arccos with the same library as the reference.
//
// Targets are returned per source in ascending target index; the caller orders targets
// by (healpix, position in the healpix list), which is the reference's neighbour order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/picca_hip.h"

namespace {

struct Grid {
    double h = 1.0;
    int64_t n = 1;
    std::vector<int64_t> key;    // sorted cell keys of the targets
    std::vector<int64_t> index;  // target index for each entry of key
};

int64_t cell_coord(double v, double h, int64_t n) {
    int64_t c = (int64_t)std::floor((v + 1.0) / h);
    return std::min<int64_t>(std::max<int64_t>(c, 0), n - 1);
}

void build_grid(Grid &g, int64_t n_tgt, const double *tx, const double *ty, const double *tz,
                double h) {
    g.h = h;
    g.n = std::max<int64_t>(1, (int64_t)std::ceil(2.0 / h) + 1);
    std::vector<std::pair<int64_t, int64_t>> items((size_t)n_tgt);
    for (int64_t t = 0; t < n_tgt; ++t) {
        int64_t cx = cell_coord(tx[t], h, g.n), cy = cell_coord(ty[t], h, g.n),
                cz = cell_coord(tz[t], h, g.n);
        items[t] = {(cx * g.n + cy) * g.n + cz, t};
    }
    std::sort(items.begin(), items.end());
    g.key.resize(items.size());
    g.index.resize(items.size());
    for (size_t k = 0; k < items.size(); ++k) {
        g.key[k] = items[k].first;
        g.index[k] = items[k].second;
    }
}

template <typename Emit>
void scan_source(const Grid &g, int64_t s, const double *sx, const double *sy, const double *sz,
                 const double *sra, const double *sdec, const int64_t *sid, const double *tx,
                 const double *ty, const double *tz, const double *tra, const double *tdec,
                 const int64_t *tid, double cos_min, double small_angle, int auto_mode,
                 std::vector<std::pair<int64_t, double>> &scratch, Emit emit) {
    scratch.clear();
    const int64_t cx = cell_coord(sx[s], g.h, g.n), cy = cell_coord(sy[s], g.h, g.n),
                  cz = cell_coord(sz[s], g.h, g.n);
    for (int64_t dx = -1; dx <= 1; ++dx) {
        const int64_t x = cx + dx;
        if (x < 0 || x >= g.n) continue;
        for (int64_t dy = -1; dy <= 1; ++dy) {
            const int64_t y = cy + dy;
            if (y < 0 || y >= g.n) continue;
            for (int64_t dz = -1; dz <= 1; ++dz) {
                const int64_t z = cz + dz;
                if (z < 0 || z >= g.n) continue;
                const int64_t key = (x * g.n + y) * g.n + z;
                auto lo = std::lower_bound(g.key.begin(), g.key.end(), key);
                for (auto it = lo; it != g.key.end() && *it == key; ++it) {
                    const int64_t t = g.index[it - g.key.begin()];
                    if (tid[t] == sid[s]) continue;
                    if (auto_mode && !(sra[s] > tra[t])) continue;
                    const double cosv = tx[t] * sx[s] + ty[t] * sy[s] + tz[t] * sz[s];
                    const bool close = std::fabs(tra[t] - sra[s]) < small_angle &&
                                       std::fabs(tdec[t] - sdec[s]) < small_angle;
                    if (cosv >= cos_min || close) scratch.emplace_back(t, cosv);
                }
            }
        }
    }
    std::sort(scratch.begin(), scratch.end(),
              [](const std::pair<int64_t, double> &a, const std::pair<int64_t, double> &b) {
                  return a.first < b.first;
              });
    emit(scratch);
}

}  // namespace

extern "C" {

// Pass 1: number of candidates per source (counts[n_src]).  Pass 2 (offsets != NULL):
// candidates written to out_target / out_cos at offsets[s] .. offsets[s] + counts[s].
// auto_mode keeps only targets with ra_source > ra_target (py/picca/cf.py:129-135).
int picca_neighbour_candidates(int64_t n_src, const double *sx, const double *sy,
                               const double *sz, const double *sra, const double *sdec,
                               const int64_t *sid, int64_t n_tgt, const double *tx,
                               const double *ty, const double *tz, const double *tra,
                               const double *tdec, const int64_t *tid, double cos_min,
                               double chord_max, double small_angle, int32_t auto_mode,
                               int64_t *counts, const int64_t *offsets, int64_t *out_target,
                               double *out_cos) {
    if (n_src < 0 || n_tgt < 0 || !counts) return PICCA_E_VALUE;
    if (!(chord_max > 0)) return PICCA_E_VALUE;
    Grid g;
    build_grid(g, n_tgt, tx, ty, tz, std::min(chord_max * (1.0 + 1e-9) + 1e-12, 4.0));
    unsigned hw = std::thread::hardware_concurrency();
    int64_t nthreads = std::max<int64_t>(1, std::min<int64_t>(hw ? hw : 1, 32));
    if (n_src < 256) nthreads = 1;
    std::vector<std::thread> pool;
    const int64_t step = (n_src + nthreads - 1) / std::max<int64_t>(nthreads, 1);
    for (int64_t th = 0; th < nthreads; ++th) {
        const int64_t lo = th * step, hi = std::min(n_src, lo + step);
        if (lo >= hi) break;
        pool.emplace_back([=, &g] {
            std::vector<std::pair<int64_t, double>> scratch;
            for (int64_t s = lo; s < hi; ++s) {
                scan_source(g, s, sx, sy, sz, sra, sdec, sid, tx, ty, tz, tra, tdec, tid,
                            cos_min, small_angle, auto_mode, scratch,
                            [&](const std::vector<std::pair<int64_t, double>> &found) {
                                if (!offsets) {
                                    counts[s] = (int64_t)found.size();
                                    return;
                                }
                                const int64_t base = offsets[s];
                                const int64_t n = std::min<int64_t>((int64_t)found.size(), counts[s]);
                                for (int64_t k = 0; k < n; ++k) {
                                    out_target[base + k] = found[k].first;
                                    out_cos[base + k] = found[k].second;
                                }
                            });
            }
        });
    }
    for (auto &t : pool) t.join();
    return PICCA_OK;
}

}  // extern "C"
