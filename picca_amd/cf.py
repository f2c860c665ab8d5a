"""Drop-in for picca's ``picca.cf`` module on the MI355X engine.

Same module globals (``py/picca/cf.py:28-79``), read at call time, and the same entry
points:

* ``fill_neighs(healpixs)`` (``py/picca/cf.py:82-135``) sets ``delta.neighbours``;
* ``compute_xi(healpixs)`` (``py/picca/cf.py:138-247``) returns
  ``(weights, xi, r_par, r_trans, z, num_pairs)`` for all pairs of the given healpixs,
  evaluated by the HIP pair kernel in one launch;
* ``compute_xi_forest_pairs_fast`` / ``compute_xi_forest_pairs`` (``py/picca/cf.py:250-387``)
  accumulate one forest pair in place into caller arrays (one launch per call; use
  ``compute_xi`` for batches);
* ``compute_dmat(healpixs)`` and ``compute_dmat_forest_pairs_fast``
  (``py/picca/cf.py:390-887``): the distortion matrix on the engine.

Additionally ``compute_xi_rows(healpixs)`` evaluates every healpix as its own output row in
a single launch, with the neighbour search done on flat arrays; ``picca_cf`` uses it.

There is no CPU path: without the HIP library every call raises ``RuntimeError``.
"""
import numpy as np

from . import _native
from . import engine as _engine
from . import metal as _metal
from .data import angle_between
from .neighbours import find_neighbours
from .store import ForestStore
from .utils import userprint

num_bins_r_par = None
num_bins_r_trans = None
num_model_bins_r_trans = None
num_model_bins_r_par = None
r_par_max = None
r_par_min = None
z_min_pairs = None
z_max_pairs = None
r_trans_max = None
ang_max = None
nside = None

zerr_cut_deg = None
zerr_cut_kms = None

counter = None
num_data = None
num_data2 = None

z_ref = None
alpha = None
alpha2 = None
alpha_abs = None
lambda_abs = None
lambda_abs2 = None

data = None
data2 = None

cosmo = None

# Wick covariance inputs (set by picca_wick): nearest-neighbour interpolations of the 1-D
# correlation and variance per delta field name, the 3-D correlations for diagrams 4-5
xi_1d = {}
get_variance_1d = {}
xi_wick = {}
max_diagram = None

reject = None
lock = None
x_correlation = False
rmu_binning = False
ang_correlation = False
remove_same_half_plate_close_pairs = False

redshift_evolution_in_distortion_matrix = True

# engine settings (no counterpart in picca)
edge_tol = _engine.DEFAULT_EDGE_TOL
device = 0
last_report = None


class _Globals:
    """Snapshot of this module's globals in the form params_from_globals expects."""

    def __init__(self):
        for name in ("num_bins_r_par", "num_bins_r_trans", "num_model_bins_r_par",
                     "num_model_bins_r_trans", "r_par_min", "r_par_max", "r_trans_max",
                     "z_min_pairs", "z_max_pairs", "zerr_cut_deg", "zerr_cut_kms", "z_ref",
                     "alpha", "alpha2", "x_correlation", "rmu_binning", "ang_correlation",
                     "remove_same_half_plate_close_pairs",
                     "redshift_evolution_in_distortion_matrix"):
            setattr(self, name, globals()[name])


def current_params():
    """picca_cf_params built from the module globals as they are now."""
    return _engine.params_from_globals(_Globals(), edge_tol=edge_tol)


# ---------------------------------------------------------------------------
# forest catalogue: data (and data2) packed once into a resident store
# ---------------------------------------------------------------------------

class _Catalogue:
    def __init__(self, data_1, data_2, angular=False):
        self.healpixs1 = sorted(data_1)
        deltas1 = [d for hp in self.healpixs1 for d in data_1[hp]]
        deltas2 = []
        if data_2 is not None:
            self.healpixs2 = sorted(data_2)
            deltas2 = [d for hp in self.healpixs2 for d in data_2[hp]]
        self.deltas = deltas1 + deltas2
        self.num1 = len(deltas1)
        self.store = ForestStore.from_deltas(self.deltas, angular=angular)
        self.index = {id(d): k for k, d in enumerate(self.deltas)}
        # targets in the reference's neighbour order: healpix ascending, then list order
        if data_2 is None:
            self.target_index = np.arange(self.num1, dtype=np.int64)
        else:
            self.target_index = np.arange(self.num1, len(self.deltas), dtype=np.int64)
        # position of each healpix's forests in the store
        self.range1 = {}
        start = 0
        for hp in self.healpixs1:
            self.range1[hp] = (start, start + len(data_1[hp]))
            start += len(data_1[hp])
        self.key = (id(data_1), id(data_2), bool(angular), len(self.deltas))

    def index_of(self, delta):
        try:
            return self.index[id(delta)]
        except KeyError:
            raise RuntimeError("delta is not part of cf.data / cf.data2") from None


_catalogue_cache = None


def _catalogue():
    global _catalogue_cache
    if data is None:
        raise RuntimeError("cf.data is not set")
    key = (id(data), id(data2), bool(ang_correlation))
    if _catalogue_cache is None or _catalogue_cache.key[:3] != key:
        _catalogue_cache = _Catalogue(data, data2, angular=bool(ang_correlation))
    return _catalogue_cache


def _same_half_plate(delta1, delta2):
    """py/picca/cf.py:171-185"""
    if remove_same_half_plate_close_pairs:
        if isinstance(delta1.fiberid, str) or isinstance(delta2.fiberid, str):
            raise RuntimeError(
                "Trying to figure out if two spectra come from the same half plate but "
                "combined reobservations were given")
        return (delta1.plate == delta2.plate) and (
            (delta1.fiberid <= 500 and delta2.fiberid <= 500)
            or (delta1.fiberid > 500 and delta2.fiberid > 500))
    return False


def _same_half_plate_flags(store, pair_a, pair_b):
    if not remove_same_half_plate_close_pairs:
        return np.zeros(len(pair_a), dtype=np.uint8)
    plate_a, plate_b = store.plate[pair_a], store.plate[pair_b]
    fib_a, fib_b = store.fiberid[pair_a], store.fiberid[pair_b]
    same = (plate_a == plate_b) & (((fib_a <= 500) & (fib_b <= 500))
                                   | ((fib_a > 500) & (fib_b > 500)))
    return same.astype(np.uint8)


def _tick():
    if counter is None or lock is None:
        return
    with lock:
        xicounter = round(counter.value * 100.0 / num_data, 2)
        if counter.value % 1000 == 0:
            userprint("computing xi: {}%".format(xicounter))
        counter.value += 1


# ---------------------------------------------------------------------------
# neighbours
# ---------------------------------------------------------------------------

def fill_neighs(healpixs):
    """Store the neighbour list of every delta of ``healpixs`` in ``delta.neighbours``."""
    cat = _catalogue()
    sources = []
    for healpix in healpixs:
        for delta in data[healpix]:
            sources.append(delta)
    if not sources:
        return
    src_index = np.array([cat.index_of(d) for d in sources], dtype=np.int64)
    nl = find_neighbours(cat.store, src_index, cat.target_index, ang_max,
                         auto=data2 is None)
    targets = [cat.deltas[k] for k in cat.target_index]
    for s, delta in enumerate(sources):
        lo, hi = nl.offsets[s], nl.offsets[s + 1]
        neigh = [targets[t] for t in nl.target[lo:hi]]
        if data2 is not None:
            delta.neighbours = np.array(neigh, dtype=object)
        else:
            delta.neighbours = neigh


# ---------------------------------------------------------------------------
# correlation function
# ---------------------------------------------------------------------------

def _run_xi(cat, pair_a, pair_b, pair_ang, pair_flags, pair_row, num_rows):
    global last_report
    params = current_params()
    eng = _engine.get_engine(device)
    eng.set_forests(cat.store, key=cat.key)
    eng.set_xi_pairs(pair_a, pair_b, pair_ang, pair_flags, pair_row, num_rows)
    result = eng.xi(params)
    last_report = result.report
    return result


def compute_xi(healpixs):
    """Correlation function of all pairs whose delta1 is in ``healpixs``."""
    cat = _catalogue()
    pair_a, pair_b, pair_ang, pair_flags = [], [], [], []
    for healpix in healpixs:
        for delta1 in data[healpix]:
            _tick()
            neighbours = delta1.neighbours
            if neighbours is None or len(neighbours) == 0:
                delta1.neighbours = None
                continue
            neighbours = list(neighbours)
            a = cat.index_of(delta1)
            b = np.array([cat.index_of(d) for d in neighbours], dtype=np.int64)
            ang = angle_between(delta1.x_cart, delta1.y_cart, delta1.z_cart, delta1.ra,
                                delta1.dec, delta1.cos_dec,
                                np.array([d.x_cart for d in neighbours]),
                                np.array([d.y_cart for d in neighbours]),
                                np.array([d.z_cart for d in neighbours]),
                                np.array([d.ra for d in neighbours]),
                                np.array([d.dec for d in neighbours]))
            flags = np.array([_same_half_plate(delta1, d) for d in neighbours],
                             dtype=np.uint8)
            pair_a.append(np.full(len(b), a, dtype=np.int64))
            pair_b.append(b)
            pair_ang.append(np.asarray(ang, dtype=np.float64))
            pair_flags.append(flags)
            delta1.neighbours = None
    nb = num_bins_r_par * num_bins_r_trans
    if not pair_a:
        return (np.zeros(nb), np.zeros(nb), np.zeros(nb), np.zeros(nb), np.zeros(nb),
                np.zeros(nb, dtype=np.int64))
    pair_a = np.concatenate(pair_a)
    result = _run_xi(cat, pair_a, np.concatenate(pair_b), np.concatenate(pair_ang),
                     np.concatenate(pair_flags), np.zeros(len(pair_a), dtype=np.int32), 1)
    we, xi, rp, rt, z, npairs = result.normalised()
    return we[0], xi[0], rp[0], rt[0], z[0], npairs[0]


def build_xi_pairs(healpixs):
    """Pair list of ``compute_xi_rows``: one output row per healpix of ``healpixs``.

    Returns (catalogue, pair_a, pair_b, pair_ang, pair_flags, pair_row).
    """
    cat = _catalogue()
    src, row_of = [], []
    for row, healpix in enumerate(healpixs):
        lo, hi = cat.range1[healpix]
        src.append(np.arange(lo, hi, dtype=np.int64))
        row_of.append(np.full(hi - lo, row, dtype=np.int32))
    src = np.concatenate(src) if src else np.zeros(0, dtype=np.int64)
    row_of = np.concatenate(row_of) if row_of else np.zeros(0, dtype=np.int32)
    nl = find_neighbours(cat.store, src, cat.target_index, ang_max, auto=data2 is None)
    s_of = nl.source_index()
    pair_a = src[s_of]
    pair_b = cat.target_index[nl.target]
    if remove_same_half_plate_close_pairs:
        # either side of a pair with a combined-reobservation id (py/picca/cf.py:171-179)
        for k in np.unique(np.concatenate([pair_a, pair_b])):
            if isinstance(cat.deltas[k].fiberid, str):
                raise RuntimeError(
                    "Trying to figure out if two spectra come from the same half plate "
                    "but combined reobservations were given")
    flags = _same_half_plate_flags(cat.store, pair_a, pair_b)
    return cat, pair_a, pair_b, nl.angle, flags, row_of[s_of]


def compute_xi_rows(healpixs):
    """``[fill_neighs([hp]) + compute_xi([hp]) for hp in healpixs]`` in one launch.

    Returns a list with one ``(weights, xi, r_par, r_trans, z, num_pairs)`` tuple per
    healpix, normalised per row like ``py/picca/cf.py:242-246``.
    """
    healpixs = list(healpixs)
    cat, pair_a, pair_b, pair_ang, flags, pair_row = build_xi_pairs(healpixs)
    result = _run_xi(cat, pair_a, pair_b, pair_ang, flags, pair_row, len(healpixs))
    we, xi, rp, rt, z, npairs = result.normalised()
    return [(we[k], xi[k], rp[k], rt[k], z[k], npairs[k]) for k in range(len(healpixs))]


def compute_xi_forest_pairs_fast(z1, r_comov1, dist_m1, weights1, delta1, z_qso_1, z2,
                                 r_comov2, dist_m2, weights2, delta2, z_qso_2, ang,
                                 same_half_plate, rebin_weight, rebin_xi, rebin_r_par,
                                 rebin_r_trans, rebin_z, rebin_num_pairs):
    """Adds one forest pair to the caller's arrays (``py/picca/cf.py:250-387``)."""
    store = ForestStore(
        np.array([0, len(z1), len(z1) + len(z2)], dtype=np.int64),
        np.concatenate([z1, z2]), np.concatenate([r_comov1, r_comov2]),
        np.concatenate([dist_m1, dist_m2]), np.concatenate([weights1, weights2]),
        np.concatenate([delta1, delta2]), np.array([z_qso_1, z_qso_2]),
        np.zeros(2), np.zeros(2), np.array([0, 1]))
    global last_report
    params = current_params()
    eng = _engine.get_engine(device)
    eng.set_forests(store)
    eng.set_xi_pairs([0], [1], [float(ang)], [1 if same_half_plate else 0], [0], 1)
    result = eng.xi(params)
    last_report = result.report
    rebin_weight += result.weights[0]
    rebin_xi += result.xi[0]
    rebin_r_par += result.r_par[0]
    rebin_r_trans += result.r_trans[0]
    rebin_z += result.z[0]
    rebin_num_pairs += result.num_pairs[0]


compute_xi_forest_pairs = compute_xi_forest_pairs_fast


# ---------------------------------------------------------------------------
# distortion matrix
# ---------------------------------------------------------------------------

def build_dmat_pairs(healpixs):
    """Kept pairs of compute_dmat for ``healpixs`` in reference order, with the
    rejection draws of ``py/picca/cf.py:444`` taken from numpy's global legacy stream.

    Returns (catalogue, pair_a, pair_b, pair_ang, pair_flags, num_pairs, num_pairs_used).
    """
    cat = _catalogue()
    pair_a, pair_b, pair_ang, pair_flags = [], [], [], []
    num_pairs = 0
    num_pairs_used = 0
    for healpix in healpixs:
        for delta1 in data[healpix]:
            _tick()
            if delta1.order is None:
                raise RuntimeError(
                    "Trying to compute the distortion matrix but order is not defined for "
                    "the deltas. Check previous warning to solve this issue")
            neighbours = delta1.neighbours
            if neighbours is None:
                neighbours = []
            keep = np.random.rand(len(neighbours)) > reject
            num_pairs += len(neighbours)
            num_pairs_used += keep.sum()
            kept = [d for d, k in zip(neighbours, keep) if k]
            if kept:
                for d in kept:
                    if d.order is None:
                        raise RuntimeError(
                            "Trying to compute the distortion matrix but order is not "
                            "defined for the deltas. Check previous warning to solve "
                            "this issue")
                flags = np.array([_same_half_plate(delta1, d) for d in kept], dtype=np.uint8)
                ang = angle_between(delta1.x_cart, delta1.y_cart, delta1.z_cart, delta1.ra,
                                    delta1.dec, delta1.cos_dec,
                                    np.array([d.x_cart for d in kept]),
                                    np.array([d.y_cart for d in kept]),
                                    np.array([d.z_cart for d in kept]),
                                    np.array([d.ra for d in kept]),
                                    np.array([d.dec for d in kept]))
                pair_a.append(np.full(len(kept), cat.index_of(delta1), dtype=np.int64))
                pair_b.append(np.array([cat.index_of(d) for d in kept], dtype=np.int64))
                pair_ang.append(np.asarray(ang, dtype=np.float64))
                pair_flags.append(flags)
            delta1.neighbours = None
    cat_arrays = [np.concatenate(x) if x else np.zeros(0) for x in
                  (pair_a, pair_b, pair_ang, pair_flags)]
    return (cat, cat_arrays[0].astype(np.int64), cat_arrays[1].astype(np.int64),
            cat_arrays[2].astype(np.float64), cat_arrays[3].astype(np.uint8), num_pairs,
            int(num_pairs_used))


def compute_dmat(healpixs):
    """Distortion matrix of the pairs of ``healpixs`` (``py/picca/cf.py:390-517``)."""
    global last_report
    cat, pair_a, pair_b, pair_ang, flags, num_pairs, num_pairs_used = build_dmat_pairs(
        healpixs)
    params = current_params()
    eng = _engine.get_engine(device)
    eng.set_forests(cat.store, key=cat.key)
    (weights_dmat, dmat, r_par_eff, r_trans_eff, z_eff, weight_eff,
     stats) = eng.dmat(params, pair_a, pair_b, pair_ang, flags)
    last_report = {"in_range_pairs": int(stats[0]), "selected_pairs": int(stats[1]),
                   "beyond_count_pass": int(stats[2]), "overflow_pairs": int(stats[3])}
    return (weights_dmat, dmat, r_par_eff, r_trans_eff, z_eff, weight_eff, num_pairs,
            num_pairs_used)


# ---------------------------------------------------------------------------
# metal distortion matrix
# ---------------------------------------------------------------------------

def build_metal_pairs(healpixs):
    """Kept pairs of compute_metal_dmat (``py/picca/cf.py:932-952``): one draw of numpy's
    global legacy stream per neighbour of each delta1, in healpix then list order; the
    neighbour lists are consumed.  Returns (catalogue, pair_a, pair_b, pair_ang,
    pair_flags, num_pairs, num_pairs_used)."""
    cat = _catalogue()
    pair_a, pair_b, pair_ang, pair_flags = [], [], [], []
    num_pairs = 0
    num_pairs_used = 0
    for healpix in healpixs:
        for delta1 in data[healpix]:
            _tick()
            neighbours = delta1.neighbours
            if neighbours is None:
                neighbours = []
            keep = np.random.rand(len(neighbours)) > reject
            num_pairs += len(neighbours)
            num_pairs_used += keep.sum()
            kept = [d for d, k in zip(neighbours, keep) if k]
            if kept:
                flags = np.array([_same_half_plate(delta1, d) for d in kept], dtype=np.uint8)
                ang = angle_between(delta1.x_cart, delta1.y_cart, delta1.z_cart, delta1.ra,
                                    delta1.dec, delta1.cos_dec,
                                    np.array([d.x_cart for d in kept]),
                                    np.array([d.y_cart for d in kept]),
                                    np.array([d.z_cart for d in kept]),
                                    np.array([d.ra for d in kept]),
                                    np.array([d.dec for d in kept]))
                pair_a.append(np.full(len(kept), cat.index_of(delta1), dtype=np.int64))
                pair_b.append(np.array([cat.index_of(d) for d in kept], dtype=np.int64))
                pair_ang.append(np.asarray(ang, dtype=np.float64))
                pair_flags.append(flags)
            delta1.neighbours = None
    arrays = [np.concatenate(x) if x else np.zeros(0) for x in
              (pair_a, pair_b, pair_ang, pair_flags)]
    return (cat, arrays[0].astype(np.int64), arrays[1].astype(np.int64),
            arrays[2].astype(np.float64), arrays[3].astype(np.uint8), num_pairs,
            int(num_pairs_used))


def compute_metal_dmat(healpixs, abs_igm1="LYA", abs_igm2="SiIII(1207)"):
    """Metal distortion matrix of the pairs of ``healpixs`` (``py/picca/cf.py:890-1232``):
    data bins from the Lya coordinates, model bins from delta1's pixels as ``abs_igm1``
    and delta2's as ``abs_igm2``, plus the swapped pass when the reference adds it
    (``:1092-1094``).  Uses the globals alpha_abs, cosmo, z_ref, lambda_abs(2).
    Returns (weights_dmat, dmat [nb, nm], r_par_eff, r_trans_eff, z_eff, weight_eff,
    num_pairs, num_pairs_used)."""
    global last_report
    from .constants import ABSORBER_IGM
    cat, pair_a, pair_b, pair_ang, flags, num_pairs, num_pairs_used = build_metal_pairs(
        healpixs)
    params = current_params()
    nb = num_bins_r_par * num_bins_r_trans
    nm = num_model_bins_r_par * num_model_bins_r_trans
    out = _metal.outputs(nb, nm)
    eng = _engine.get_engine(device)
    eng.set_forests(cat.store, key=cat.key)
    alpha1, alpha2_ = alpha_abs[abs_igm1], alpha_abs[abs_igm2]
    eng.metal_set_absorber(0, *_metal.absorber_table(cat.store, ABSORBER_IGM[abs_igm1],
                                                     alpha1, cosmo))
    slot2 = 0
    if abs_igm2 != abs_igm1:
        slot2 = 1
        eng.metal_set_absorber(1, *_metal.absorber_table(cat.store, ABSORBER_IGM[abs_igm2],
                                                         alpha2_, cosmo))
    den = (1 + z_ref) ** (alpha1 + alpha2_ - 2)
    mode = _native.PICCA_METAL_CROSS if x_correlation else _native.PICCA_METAL_AUTO
    reports = [eng.metal_pass(params, mode, 0, slot2, den, pair_a, pair_b, pair_ang, flags,
                              out)]
    if ((not x_correlation) and (abs_igm1 != abs_igm2)) or (
            x_correlation and (lambda_abs == lambda_abs2)):
        reports.append(eng.metal_pass(params, mode, slot2, 0, den, pair_a, pair_b, pair_ang,
                                      flags, out))
    last_report = {"passes": reports}
    return (out[0], out[1].reshape(nb, nm), out[2], out[3], out[4], out[5], num_pairs,
            num_pairs_used)


# ---------------------------------------------------------------------------
# Wick covariance
# ---------------------------------------------------------------------------

def _weighted_xi_1d(delta):
    """(w w^T) xi_1d(|log_lambda - log_lambda^T|) sqrt(v v^T) of one delta, with the
    interpolations of its field (``py/picca/cf.py:1392-1398``)."""
    variance = get_variance_1d[delta.fname](delta.log_lambda)
    weights = delta.weights
    return (weights * weights[:, None]) * xi_1d[delta.fname](
        abs(delta.log_lambda - delta.log_lambda[:, None])) * np.sqrt(variance * variance[:, None])


def compute_wick_terms(healpixs):
    """Wick expansion of the covariance of the auto-correlation (``py/picca/cf.py:1328-1483``)
    for diagrams 1-3: one draw of numpy's global legacy stream per delta of each healpix,
    every neighbour of a kept delta, T1-T3 of all pairs in one GPU computation.  Returns
    (weights_wick, num_pairs_wick, num_pairs, num_pairs_used, t1, t2, t3, t4, t5, t6);
    t4-t6 are zero (diagrams 4-5 need max_diagram > 3, which this engine refuses)."""
    global last_report
    if max_diagram is not None and max_diagram > 3:
        raise NotImplementedError("diagrams 4 and 5 (max_diagram > 3) are not part of this "
                                  "engine")
    cat = _catalogue()
    nb = num_bins_r_par * num_bins_r_trans
    pair_a, pair_b, pair_ang = [], [], []
    num_pairs = 0
    num_pairs_used = 0
    for healpix in healpixs:
        w = np.random.rand(len(data[healpix])) > reject
        num_pairs += len(data[healpix])
        num_pairs_used += w.sum()
        if w.sum() == 0:
            continue
        for delta1 in [delta for index, delta in enumerate(data[healpix]) if w[index]]:
            _tick()
            neighbours = delta1.neighbours
            if neighbours is None or len(neighbours) == 0:
                continue
            ang = angle_between(delta1.x_cart, delta1.y_cart, delta1.z_cart, delta1.ra,
                                delta1.dec, delta1.cos_dec,
                                np.array([d.x_cart for d in neighbours]),
                                np.array([d.y_cart for d in neighbours]),
                                np.array([d.z_cart for d in neighbours]),
                                np.array([d.ra for d in neighbours]),
                                np.array([d.dec for d in neighbours]))
            pair_a.append(np.full(len(neighbours), cat.index_of(delta1), dtype=np.int64))
            pair_b.append(np.array([cat.index_of(d) for d in neighbours], dtype=np.int64))
            pair_ang.append(np.asarray(ang, dtype=np.float64))
    pair_a, pair_b, pair_ang = [np.concatenate(x) if x else np.zeros(0) for x in
                                (pair_a, pair_b, pair_ang)]
    store = cat.store
    wx_offsets = np.full(store.num_forests, -1, dtype=np.int64)
    mats = []
    used = 0
    for f in np.unique(np.concatenate([pair_a, pair_b]).astype(np.int64)):
        m = _weighted_xi_1d(cat.deltas[f]).ravel()
        wx_offsets[f] = used
        used += len(m)
        mats.append(m)
    wx = np.concatenate(mats) if mats else np.zeros(0)
    zev1 = ((1 + store.z) / (1 + z_ref)) ** (alpha - 1)
    zev2 = ((1 + store.z) / (1 + z_ref)) ** (alpha2 - 1)
    eng = _engine.get_engine(device)
    eng.set_forests(store, key=cat.key)
    weights_wick, num_pairs_wick, t1, t2, t3, report = eng.wick_t123(
        current_params(), pair_a, pair_b, pair_ang, zev1, zev2, wx_offsets, wx)
    last_report = report
    zero = np.zeros((nb, nb))
    return (weights_wick, num_pairs_wick, num_pairs, num_pairs_used, t1, t2, t3, zero,
            zero.copy(), zero.copy())


def compute_dmat_forest_pairs_fast(log_lambda1, log_lambda2, r_comov1, r_comov2, dist_m1,
                                   dist_m2, z1, z2, weights1, weights2, z_qso_1, z_qso_2,
                                   ang, weights_dmat, dmat, r_par_eff, r_trans_eff, z_eff,
                                   weight_eff, same_half_plate, order1, order2):
    """Adds one forest pair to the caller's dmat arrays (``py/picca/cf.py:520-887``)."""
    store = ForestStore(
        np.array([0, len(z1), len(z1) + len(z2)], dtype=np.int64),
        np.concatenate([z1, z2]), np.concatenate([r_comov1, r_comov2]),
        np.concatenate([dist_m1, dist_m2]), np.concatenate([weights1, weights2]),
        np.zeros(len(z1) + len(z2)), np.array([z_qso_1, z_qso_2]), np.zeros(2), np.zeros(2),
        np.array([0, 1]), log_lambda=np.concatenate([log_lambda1, log_lambda2]),
        order=np.array([order1, order2]))
    params = current_params()
    eng = _engine.get_engine(device)
    eng.set_forests(store)
    out = eng.dmat(params, [0], [1], [float(ang)], [1 if same_half_plate else 0])
    weights_dmat += out[0]
    dmat += out[1].reshape(dmat.shape)
    r_par_eff += out[2]
    r_trans_eff += out[3]
    z_eff += out[4]
    weight_eff += out[5]
