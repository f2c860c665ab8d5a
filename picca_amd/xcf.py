"""Forest x quasar cross-correlation on the MI355X engine: mirror of ``picca.xcf``.

Module globals with picca's names (py/picca/xcf.py:20-67), and:

* ``fill_neighs`` (:71-128): for every delta, the quasars of the healpix pixels near it
  (healpix ascending, catalogue order) minus the same thingid, minus angle >= ang_max,
  minus the quasar pairs cut by the redshift-error cut, minus (outside the angular mode)
  those whose r_comov puts every pixel out of the r_par range;
* ``compute_xi`` (:131-220): the pair kernel over every (delta, neighbour) of the given
  healpix pixels in one GPU launch, normalised like the reference;
* ``compute_xi_rows``: one output row per healpix in one launch (what picca_xcf needs);
* ``compute_xi_forest_pairs_fast``: the in-place per-forest kernel (:223-322), as a
  launch of one forest;
* ``compute_dmat`` (:324-424): the distortion matrix of the kept (delta, quasar) pairs
  of the given healpix pixels in one batched GPU computation, and its per-forest kernel
  ``compute_dmat_forest_pairs_fast`` (:427-658).  Each quasar enters the engine's forest
  store as a one-pixel forest with order -1 (not projected), so the distortion-matrix
  kernels of the auto-correlation keep only the terms of the delta's continuum fit;
* ``compute_metal_dmat`` (:677-800): the metal distortion matrix of one absorber, through
  the engine's metal pass over the same store.

The neighbour candidates come from the engine's grid search; the angles are
``QSO.get_angle_between`` evaluated with numpy exactly as the reference does.
"""
import sys

import numpy as np

from . import engine as _engine
from . import _native
from . import metal as _metal
from .constants import SPEED_LIGHT
from .neighbours import find_neighbours
from .store import ForestStore

num_bins_r_par = None
num_bins_r_trans = None
num_model_bins_r_trans = None
num_model_bins_r_par = None
r_par_max = None
r_par_min = None
r_trans_max = None
z_cut_max = None
z_cut_min = None
z_min_pairs = None
z_max_pairs = None
ang_max = None
nside = None
zerr_cut_deg = None
zerr_cut_kms = None
counter = None
num_data = None
z_ref = None
alpha = None
alpha_obj = None
lambda_abs = None
alpha_abs = None
cosmo = None
data = None
objs = None
reject = None
lock = None
x_correlation = True
ang_correlation = False
rmu_binning = False
redshift_evolution_in_distortion_matrix = True

device = 0
last_report = None


class _Objects:
    """The quasar catalogue as arrays, in the reference's neighbour order."""

    def __init__(self, objs_dict):
        self.healpixs = sorted(objs_dict)
        self.objects = [o for hp in self.healpixs for o in objs_dict[hp]]

        def arr(name, default=np.nan):
            return np.array([getattr(o, name, default) if getattr(o, name, None) is not None
                             else default for o in self.objects], dtype=np.float64)
        self.x_cart = arr("x_cart")
        self.y_cart = arr("y_cart")
        self.z_cart = arr("z_cart")
        self.ra = arr("ra")
        self.dec = arr("dec")
        self.cos_dec = arr("cos_dec")
        self.thingid = np.array([int(o.thingid) for o in self.objects], dtype=np.int64)
        self.z_qso = arr("z_qso")
        self.r_comov = arr("r_comov")
        self.dist_m = arr("dist_m")
        self.weights = arr("weights")
        self.log_lambda = arr("log_lambda")
        self.key = (id(objs_dict), len(self.objects))


class _Catalogue:
    def __init__(self, data_dict, objs_dict, angular):
        self.healpixs = sorted(data_dict)
        self.deltas = [d for hp in self.healpixs for d in data_dict[hp]]
        self.store = ForestStore.from_deltas(self.deltas, angular=angular)
        self.index = {id(d): k for k, d in enumerate(self.deltas)}
        self.objects = _Objects(objs_dict)
        self.key = (id(data_dict), id(objs_dict), bool(angular), len(self.deltas))

    def index_of(self, delta):
        try:
            return self.index[id(delta)]
        except KeyError:
            raise RuntimeError("delta is not part of xcf.data") from None


_catalogue_cache = None


def _catalogue():
    global _catalogue_cache
    if data is None or objs is None:
        raise RuntimeError("xcf.data and xcf.objs must be set")
    key = (id(data), id(objs), bool(ang_correlation))
    if _catalogue_cache is None or _catalogue_cache.key[:3] != key:
        _catalogue_cache = _Catalogue(data, objs, bool(ang_correlation))
    return _catalogue_cache


def current_params():
    return _engine.params_from_globals(sys.modules[__name__])


def dmat_params():
    """The globals as distortion-matrix parameters: signed r_par, and the quasars'
    redshift-evolution exponent alpha_obj in the second-forest slot."""
    params = current_params()
    params.x_correlation = 1
    params.alpha2 = float(alpha_obj) if alpha_obj is not None else float("nan")
    return params


def _quasar_forests(store, obj):
    """The deltas' store followed by one one-pixel forest per quasar (order -1)."""
    nf = len(store.offsets) - 1
    nq = len(obj.z_qso)
    offsets = np.concatenate([store.offsets,
                              store.offsets[-1] + np.arange(1, nq + 1, dtype=np.int64)])
    ll = store.log_lambda if store.log_lambda is not None else np.zeros(len(store.z))
    return ForestStore(
        offsets, np.concatenate([store.z, obj.z_qso]),
        np.concatenate([store.r_comov, obj.r_comov]),
        np.concatenate([store.dist_m, obj.dist_m]),
        np.concatenate([store.weights, obj.weights]),
        np.concatenate([store.delta, np.zeros(nq)]),
        np.concatenate([store.z_qso, obj.z_qso]),
        np.concatenate([store.ra, obj.ra]), np.concatenate([store.dec, obj.dec]),
        np.concatenate([store.thingid, obj.thingid]),
        log_lambda=np.concatenate([ll, np.zeros(nq)]),
        order=np.concatenate([store.order, np.full(nq, -1, dtype=np.int32)])), nf


def fill_neighs(healpixs):
    """Store the quasar neighbours of every delta of ``healpixs`` in ``delta.neighbours``
    (an object array) and their angles in ``delta.neighbour_angles``."""
    cat = _catalogue()
    sources = [d for hp in healpixs for d in data[hp]]
    if not sources:
        return
    src_index = np.array([cat.index_of(d) for d in sources], dtype=np.int64)
    obj = cat.objects
    nl = find_neighbours(cat.store, src_index, np.arange(len(obj.objects), dtype=np.int64),
                         ang_max, auto=False, target_store=obj)
    for s, delta in enumerate(sources):
        lo, hi = nl.offsets[s], nl.offsets[s + 1]
        tgt = nl.target[lo:hi]
        ang = nl.angle[lo:hi]
        w = np.ones(len(tgt), dtype=bool)
        if zerr_cut_deg is not None:
            ang_deg = 180.0 / np.pi * ang
            obj_z_qso = obj.z_qso[tgt]
            z_qq = 0.5 * (delta.z_qso + obj_z_qso)
            dv_kms = np.abs(delta.z_qso - obj_z_qso) / (1 + z_qq)
            dv_kms *= SPEED_LIGHT
            zerr_cut_mask = ang_deg < zerr_cut_deg
            zerr_cut_mask &= dv_kms < zerr_cut_kms
            w &= ~zerr_cut_mask
        if not ang_correlation:
            r_comov = obj.r_comov[tgt]
            f = r_trans_max if rmu_binning else 1
            w &= (delta.r_comov[0] - r_comov) * np.cos(ang / 2.0) < r_par_max * f
            w &= (delta.r_comov[-1] - r_comov) * np.cos(ang / 2.0) > r_par_min * f
        delta.neighbours = np.array([obj.objects[t] for t in tgt[w]], dtype=object)
        delta.neighbour_index = tgt[w]
        delta.neighbour_angles = ang[w]


def _pairs(cat, deltas_by_row):
    """Flat pair arrays for a list of (row, deltas) in row order."""
    obj = cat.objects
    forest, oidx, ang, row = [], [], [], []
    for r, deltas in deltas_by_row:
        for delta in deltas:
            idx = getattr(delta, "neighbour_index", None)
            if idx is None or len(idx) == 0:
                continue
            forest.append(np.full(len(idx), cat.index_of(delta), dtype=np.int32))
            oidx.append(np.asarray(idx, dtype=np.int64))
            ang.append(np.asarray(delta.neighbour_angles, dtype=np.float64))
            row.append(np.full(len(idx), r, dtype=np.int32))
    if not forest:
        return None
    forest = np.concatenate(forest)
    oidx = np.concatenate(oidx)
    ang = np.concatenate(ang)
    row = np.concatenate(row)
    if ang_correlation:
        r2 = 10.0 ** obj.log_lambda[oidx]
        d2 = r2
    else:
        r2 = obj.r_comov[oidx]
        d2 = obj.dist_m[oidx]
    return forest, r2, d2, obj.z_qso[oidx], obj.weights[oidx], ang, row


def _run(cat, pairs, num_rows):
    global last_report
    params = current_params()
    eng = _engine.get_engine(device)
    eng.set_forests(cat.store, key=cat.key)
    if pairs is None:
        nb = params.num_bins_r_par * params.num_bins_r_trans
        zeros = [np.zeros((num_rows, nb)) for _ in range(5)]
        last_report = None
        return zeros + [np.zeros((num_rows, nb), dtype=np.int64)]
    forest, r2, d2, z2, w2, ang, row = pairs
    eng.xcf_set_pairs(forest, r2, d2, z2, w2, ang, row, num_rows, angular=bool(ang_correlation))
    res = eng.xcf(params)
    last_report = res.report
    return [res.weights, res.xi, res.r_par, res.r_trans, res.z, res.num_pairs]


def _normalise(we, xi, rp, rt, z):
    w = we > 0
    for arr in (xi, rp, rt, z):
        arr[w] /= we[w]


def compute_xi(healpixs):
    """compute_xi (py/picca/xcf.py:131-220) over the deltas of ``healpixs``; consumes
    their neighbour lists."""
    cat = _catalogue()
    deltas = [d for hp in healpixs for d in data[hp]]
    out = _run(cat, _pairs(cat, [(0, deltas)]), 1)
    we, xi, rp, rt, z = (a[0].copy() for a in out[:5])
    num_pairs = out[5][0].copy()
    _normalise(we, xi, rp, rt, z)
    for d in deltas:
        d.neighbours = None
    return we, xi, rp, rt, z, num_pairs


def compute_xi_rows(healpixs):
    """One normalised compute_xi row per healpix, all in one launch (fills the neighbours
    first)."""
    fill_neighs(healpixs)
    cat = _catalogue()
    groups = [(r, data[hp]) for r, hp in enumerate(healpixs)]
    out = _run(cat, _pairs(cat, groups), len(healpixs))
    rows = []
    for r in range(len(healpixs)):
        we, xi, rp, rt, z = (a[r].copy() for a in out[:5])
        _normalise(we, xi, rp, rt, z)
        rows.append((we, xi, rp, rt, z, out[5][r].copy()))
    for hp in healpixs:
        for d in data[hp]:
            d.neighbours = None
    return rows


def compute_xi_forest_pairs_fast(z1, r_comov1, dist_m1, weights1, delta1, z2, r_comov2,
                                 dist_m2, weights2, ang, rebin_weight, rebin_xi, rebin_r_par,
                                 rebin_r_trans, rebin_z, rebin_num_pairs):
    """Adds one forest against its quasars to the caller's arrays (py/picca/xcf.py:223-322);
    runs as a GPU launch of one forest."""
    n1 = len(z1)
    store = ForestStore(np.array([0, n1], dtype=np.int64), z1, r_comov1, dist_m1, weights1,
                        delta1, np.zeros(1), np.zeros(1), np.zeros(1), np.array([0]))
    params = current_params()
    eng = _engine.get_engine(device)
    eng.set_forests(store)
    n2 = len(z2)
    eng.xcf_set_pairs(np.zeros(n2, np.int32), r_comov2, dist_m2, z2, weights2, ang,
                      np.zeros(n2, np.int32), 1, angular=bool(ang_correlation))
    res = eng.xcf(params)
    rebin_weight += res.weights[0]
    rebin_xi += res.xi[0]
    rebin_r_par += res.r_par[0]
    rebin_r_trans += res.r_trans[0]
    rebin_z += res.z[0]
    rebin_num_pairs += res.num_pairs[0]


_dmat_store_cache = None


def _dmat_store(cat):
    global _dmat_store_cache
    if _dmat_store_cache is None or _dmat_store_cache[0] != cat.key:
        store, nf = _quasar_forests(cat.store, cat.objects)
        _dmat_store_cache = (cat.key, store, nf)
    return _dmat_store_cache[1], _dmat_store_cache[2]


def build_dmat_pairs(healpixs):
    """Kept (delta, quasar) pairs of compute_dmat (py/picca/xcf.py:357-407) in reference
    order: one draw of numpy's global legacy stream per quasar neighbour of each delta; a
    delta that keeps no quasar is skipped before its pairs are counted and keeps its
    neighbour list.  Returns (forest index, quasar index, angle, num_pairs,
    num_pairs_used)."""
    cat = _catalogue()
    fa, qb, angs = [], [], []
    num_pairs = 0
    num_pairs_used = 0
    for healpix in healpixs:
        for delta1 in data[healpix]:
            if delta1.order is None:
                raise RuntimeError(
                    "Trying to compute the distortion matrix but order is not defined for "
                    "the deltas. Check previous warning to solve this issue")
            idx = getattr(delta1, "neighbour_index", None)
            n = 0 if idx is None or delta1.neighbours is None else len(idx)
            keep = np.random.rand(n) > reject
            if keep.sum() == 0:
                continue
            num_pairs += n
            num_pairs_used += int(keep.sum())
            fa.append(np.full(int(keep.sum()), cat.index_of(delta1), dtype=np.int64))
            qb.append(np.asarray(idx, dtype=np.int64)[keep])
            angs.append(np.asarray(delta1.neighbour_angles, dtype=np.float64)[keep])
            delta1.neighbours = None
    cat_arrays = [np.concatenate(x) if x else np.zeros(0) for x in (fa, qb, angs)]
    return (cat_arrays[0].astype(np.int64), cat_arrays[1].astype(np.int64),
            cat_arrays[2].astype(np.float64), num_pairs, num_pairs_used)


def compute_dmat(healpixs):
    """compute_dmat (py/picca/xcf.py:324-424) over the deltas of ``healpixs``; consumes
    their neighbour lists (fill_neighs first).  Returns (weights_dmat, dmat [nb, nm],
    r_par_eff, r_trans_eff, z_eff, weight_eff, num_pairs, num_pairs_used)."""
    global last_report
    cat = _catalogue()
    fa, qb, ang, num_pairs, num_pairs_used = build_dmat_pairs(healpixs)
    store, nf = _dmat_store(cat)
    eng = _engine.get_engine(device)
    eng.set_forests(store, key=("xdmat",) + cat.key)
    flags = np.full(len(fa), _native.PICCA_PAIR_XCF_COUNT, dtype=np.uint8)
    (weights_dmat, dmat, r_par_eff, r_trans_eff, z_eff, weight_eff,
     stats) = eng.dmat(dmat_params(), fa, nf + qb, ang, flags)
    last_report = {"in_range_pairs": int(stats[0]), "selected_pairs": int(stats[1]),
                   "beyond_count_pass": int(stats[2]), "overflow_pairs": int(stats[3])}
    return (weights_dmat, dmat, r_par_eff, r_trans_eff, z_eff, weight_eff, num_pairs,
            num_pairs_used)


def compute_dmat_forest_pairs_fast(log_lambda1, r_comov1, dist_m1, z1, weights1, r_comov2,
                                   dist_m2, z2, weights2, ang, weights_dmat, dmat, r_par_eff,
                                   r_trans_eff, z_eff, weight_eff, order1):
    """Adds one delta against its kept quasars to the caller's arrays
    (py/picca/xcf.py:427-658); runs as one GPU computation."""
    n1 = len(z1)
    n2 = len(z2)
    offsets = np.concatenate([[0, n1], n1 + np.arange(1, n2 + 1)]).astype(np.int64)
    store = ForestStore(
        offsets, np.concatenate([z1, z2]), np.concatenate([r_comov1, r_comov2]),
        np.concatenate([dist_m1, dist_m2]), np.concatenate([weights1, weights2]),
        np.zeros(n1 + n2), np.concatenate([[0.0], z2]), np.zeros(n2 + 1), np.zeros(n2 + 1),
        np.arange(n2 + 1), log_lambda=np.concatenate([log_lambda1, np.zeros(n2)]),
        order=np.concatenate([[order1], np.full(n2, -1)]))
    eng = _engine.get_engine(device)
    eng.set_forests(store)
    out = eng.dmat(dmat_params(), np.zeros(n2, np.int64), 1 + np.arange(n2), ang,
                   np.full(n2, _native.PICCA_PAIR_XCF_COUNT, dtype=np.uint8))
    weights_dmat += out[0]
    dmat += out[1].reshape(dmat.shape)
    r_par_eff += out[2]
    r_trans_eff += out[3]
    z_eff += out[4]
    weight_eff += out[5]


def compute_metal_dmat(healpixs, abs_igm="SiII(1526)"):
    """Metal distortion matrix of the deltas of ``healpixs`` against their quasar
    neighbours (``py/picca/xcf.py:677-800``) for absorber ``abs_igm`` on the forest side:
    a delta none of whose pixels lies below its quasar's redshift as ``abs_igm`` is skipped
    before its draw; then one draw of numpy's global legacy stream per quasar neighbour.
    Uses the globals alpha_abs, cosmo, z_ref.  Returns (weights_dmat, dmat [nb, nm],
    r_par_eff, r_trans_eff, z_eff, weight_eff, num_pairs, num_pairs_used)."""
    global last_report
    from .constants import ABSORBER_IGM
    cat = _catalogue()
    store, nf = _dmat_store(cat)
    lam = ABSORBER_IGM[abs_igm]
    forest_tab = _metal.absorber_table(cat.store, lam, alpha_abs[abs_igm], cosmo,
                                       z_ref=z_ref)
    obj = cat.objects
    qso_tab = _metal.quasar_table(obj.z_qso, obj.r_comov, obj.dist_m, obj.weights)
    table = [np.concatenate([x, y]) for x, y in zip(forest_tab, qso_tab)]
    # pixels of each delta whose absorber redshift lies below its quasar's
    offsets = cat.store.offsets
    mask = forest_tab[0] < np.repeat(cat.store.z_qso, np.diff(offsets))
    cum = np.concatenate([[0], np.cumsum(mask, dtype=np.int64)])
    below = cum[offsets[1:]] - cum[offsets[:-1]]
    fa, qb, angs = [], [], []
    num_pairs = 0
    num_pairs_used = 0
    for healpix in healpixs:
        for delta1 in data[healpix]:
            k = cat.index_of(delta1)
            if below[k] == 0:
                continue
            idx = getattr(delta1, "neighbour_index", None)
            n = 0 if idx is None or delta1.neighbours is None else len(idx)
            keep = np.random.rand(n) > reject
            num_pairs += n
            num_pairs_used += int(keep.sum())
            if keep.any():
                fa.append(np.full(int(keep.sum()), k, dtype=np.int64))
                qb.append(np.asarray(idx, dtype=np.int64)[keep])
                angs.append(np.asarray(delta1.neighbour_angles, dtype=np.float64)[keep])
            delta1.neighbours = None
    fa, qb, ang = [np.concatenate(x) if x else np.zeros(0) for x in (fa, qb, angs)]
    nb = num_bins_r_par * num_bins_r_trans
    nm = num_model_bins_r_par * num_model_bins_r_trans
    out = _metal.outputs(nb, nm)
    eng = _engine.get_engine(device)
    eng.set_forests(store, key=("xdmat",) + cat.key)
    eng.metal_set_absorber(0, *table)
    params = dmat_params()
    report = eng.metal_pass(params, _native.PICCA_METAL_XCF, 0, 0, 1.0,
                            fa.astype(np.int64), (nf + qb).astype(np.int64),
                            ang.astype(np.float64), np.zeros(len(fa), dtype=np.uint8), out)
    last_report = {"passes": [report]}
    return (out[0], out[1].reshape(nb, nm), out[2], out[3], out[4], out[5], num_pairs,
            num_pairs_used)
