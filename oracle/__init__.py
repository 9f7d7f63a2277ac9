"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The engine in
bmagwa_amd/ never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same sources (the sanitizer build
# of `make -C oracle asan`)
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

EF = {"A": 0, "H": 1, "D": 2, "R": 3, "AH": 4}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class Opts(C.Structure):
    """Mirror of orc_opts (the INI keys of src/options.hpp:138-321)."""
    _fields_ = [
        ("n", C.c_uint32), ("m_e", C.c_uint32), ("m_g", C.c_uint64),
        ("do_n_iter", C.c_uint64),
        ("n_rao", C.c_uint32), ("n_rao_burnin", C.c_uint32), ("thin", C.c_uint32),
        ("n_sample_tau2_and_missing", C.c_uint32),
        ("max_move_size", C.c_uint32), ("max_SNP_neighborhood_size", C.c_uint32),
        ("delay_rejection", C.c_uint32), ("seed", C.c_uint32),
        ("p_move_size", C.c_double), ("p_move_size_nbs", C.c_double),
        ("p_move_size_nbc", C.c_double), ("p_move_size_acpt_goal", C.c_double),
        ("adapt_p_move_size", C.c_int32), ("adaptation", C.c_int32),
        ("flat_proposal_dist", C.c_int32), ("save_beta", C.c_int32),
        ("sampler_type", C.c_int32),
        ("n_types", C.c_int32), ("types", C.c_int32 * 5),
        ("e_qg", C.c_double), ("var_qg", C.c_double), ("nu_sigma2", C.c_double),
        ("R2mode_sigma2", C.c_double), ("s2_sigma2", C.c_double), ("mu_alpha", C.c_double),
        ("inv_tau2_e_const_val", C.c_double), ("inv_tau2_e_val", C.c_double),
        ("use_individual_tau2", C.c_int32),
        ("types_prior", C.c_double * 5),
        ("nu_tau2", C.c_double * 4), ("eh_tau2", C.c_double * 4), ("s2_tau2", C.c_double * 4),
        ("recode_g_to_minor_allele_count", C.c_int32),
        ("chain", C.c_uint32),
    ]


def make_opts(cfg):
    """cfg: dict of INI-style keys (defaults of src/options.hpp:138-221)."""
    o = Opts()
    d = dict(max_move_size=20, max_SNP_neighborhood_size=10, p_move_size=0.2,
             p_move_size_nbs=0.2, p_move_size_nbc=0.2, adapt_p_move_size=1,
             p_move_size_acpt_goal=0.0, adaptation=0, flat_proposal_dist=0,
             save_beta=0, delay_rejection=0, R2mode_sigma2=0.0, s2_sigma2=0.0,
             mu_alpha=0.0, inv_tau2_e_val=0.0, use_individual_tau2=0,
             recode_g_to_minor_allele_count=0, m_e=0, chain=0, sampler_type=0)
    d.update(cfg)
    types = d.pop("types")
    if isinstance(types, str):
        types = [t.strip() for t in types.split(",")]
    tv = sorted(EF[t] for t in types)
    o.n_types = len(tv)
    for i, t in enumerate(tv):
        o.types[i] = t
    tp = d.pop("types_prior", [1.0] * 5)
    for i in range(5):
        o.types_prior[i] = tp[i]
    for arr in ("nu_tau2", "eh_tau2", "s2_tau2"):
        v = d.pop(arr, [0.0] * 4)
        for i in range(4):
            getattr(o, arr)[i] = v[i]
    for k, v in d.items():
        setattr(o, k, v)
    return o


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_U32 = C.POINTER(C.c_uint32)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)


class Rng(C.Structure):
    """orc_rng: Philox4x32-10 key and counter of one stream."""
    _fields_ = [("key", C.c_uint32 * 2), ("ctr", C.c_uint32 * 4)]


def mvnormal_invchol(seed, mu, U, scale, count):
    """count draws of x = scale U^-1 z + mu (src/rand.hpp:144-168) from one
    stream; U upper triangular (d x d)."""
    r = Rng()
    lib().orc_rng_init(C.byref(r), seed, 0, 0, 0)
    d = len(mu)
    Uf = np.asfortranarray(U, dtype=np.float64)
    m = np.ascontiguousarray(mu, dtype=np.float64)
    out = np.zeros((count, d))
    x = np.zeros(d)
    for i in range(count):
        lib().orc_rng_mvnormal_invchol(C.byref(r), ptr(m), Uf.ctypes.data_as(C.POINTER(C.c_double)), d, d,
                                       scale, ptr(x))
        out[i] = x
    return out


def adapt_jd_mb(p_ms, r_ms, r_n, q_p0):
    """One move-size adaptation step (src/sampler.hpp:1728-1769); q_p0 is
    updated in place."""
    a = np.ascontiguousarray(r_ms, dtype=np.float64)
    b = np.ascontiguousarray(r_n, dtype=np.float64)
    assert q_p0.dtype == np.float64 and q_p0.flags.c_contiguous
    return lib().orc_adapt_jd_mb(p_ms, ptr(a), ptr(b), ptr(q_p0), len(a))


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    _sig(lib, "orc_philox4x32_10", None, _U32, _U32, _U32)
    _sig(lib, "orc_rng_fill", None, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
         C.c_int, C.c_double, C.c_double, C.c_int64, _D)
    _sig(lib, "orc_get_genotype", C.c_double, _U8, C.c_size_t, C.c_size_t, C.c_size_t)
    _sig(lib, "orc_get_genotypes", None, _U8, C.c_size_t, C.c_uint32, C.c_int, C.c_size_t, _D)
    _sig(lib, "orc_data_create", _P, C.c_uint32, C.c_uint64, C.c_uint32, _U8, _D, _D, C.c_int)
    _sig(lib, "orc_data_destroy", None, _P)
    _sig(lib, "orc_read_bed", C.c_int, C.c_char_p, C.c_uint32, C.c_uint64, _U8)
    _sig(lib, "orc_data_body", _U8, _P)
    _sig(lib, "orc_data_stats", None, _P, _D)
    _sig(lib, "orc_data_missing", C.c_int, _P, C.c_size_t, C.POINTER(C.c_size_t), C.c_int, _D)
    _sig(lib, "orc_datamodel_create", _P, _P, _I32, C.c_int)
    _sig(lib, "orc_datamodel_destroy", None, _P)
    _sig(lib, "orc_datamodel_get", None, _P, C.c_int, C.c_size_t, _D)
    _sig(lib, "orc_datamodel_set_missing", None, _P, C.c_size_t, C.c_int, C.c_int)
    _sig(lib, "orc_datamodel_offset", C.c_int, _P, C.POINTER(C.c_int))
    _sig(lib, "orc_datamodel_update_prexx", None, _P, C.c_size_t, _D)
    _sig(lib, "orc_self_moments", None, _P, _D)
    _sig(lib, "orc_band_xtx", None, _U8, C.c_size_t, C.c_uint32, C.c_uint64, C.c_uint32,
         C.c_int, C.c_int, _I64)
    _sig(lib, "orc_chol", C.c_int, _D, C.c_int, C.c_int)
    _sig(lib, "orc_chol_update", C.c_int, _D, C.c_int, C.c_int, _D, C.c_double)
    _sig(lib, "orc_chol_downdate", None, _D, C.c_int, C.c_int, C.c_int)
    _sig(lib, "orc_chol_swapadj", None, _D, C.c_int, C.c_int, C.c_int, _D)
    _sig(lib, "orc_drotg", None, _D, _D, _D, _D)
    _sig(lib, "orc_sampler_create", _P, C.POINTER(Opts), _P)
    _sig(lib, "orc_sampler_destroy", None, _P)
    _sig(lib, "orc_sampler_init_p_proposal_flat", None, _P)
    _sig(lib, "orc_sampler_sample", C.c_int, _P, C.c_uint64)
    _sig(lib, "orc_sampler_setup", C.c_int, _P)
    _sig(lib, "orc_sampler_set_learned", C.c_int, _P, C.POINTER(C.c_float), C.c_int, C.c_double)
    _sig(lib, "orc_sampler_set_move_mix", C.c_int, _P, _D)
    _F = C.POINTER(C.c_float)
    _sig(lib, "orc_bf16_round", C.c_float, C.c_float)
    _sig(lib, "orc_mlp_forward", None, _F, _F, C.c_int64, _F)
    _sig(lib, "orc_mlp_gradients", C.c_double, _F, _F, _F, C.c_int64, _D)
    _sig(lib, "orc_mlp_loss_exact", C.c_double, _F, _F, _F, C.c_int64, _D)
    _sig(lib, "orc_mlp_adam", None, _F, _F, _F, _F, C.c_float, C.c_int)
    _sig(lib, "orc_learn_pair_features", None, _D, C.c_int, C.c_int, C.c_uint32,
         C.POINTER(C.c_int32), C.c_int, _F, _F, C.c_uint64, C.c_int, _F)
    _sig(lib, "orc_learn_snp_features", None, _D, C.c_int, C.c_int, C.c_uint32, C.c_uint64,
         C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_int32), _D,
         C.c_uint64, C.POINTER(C.c_float))
    _sig(lib, "orc_sampler_marginal_z", None, _P, _F)
    _sig(lib, "orc_sampler_swap_labels", C.c_int64, _P, C.c_int, C.POINTER(C.c_uint32),
         C.POINTER(C.c_int32), _F, C.c_int64)
    _sig(lib, "orc_sampler_iterate", C.c_int, _P, C.c_uint64)
    _sig(lib, "orc_sampler_p_rao", _D, _P)
    _sig(lib, "orc_sampler_p_proposal", _D, _P)
    _sig(lib, "orc_sampler_set_p_proposal", C.c_int, _P, _D)
    _sig(lib, "orc_sampler_trace", C.c_int64, _P, _U8, _U8, _U8, _U32, _D)
    _sig(lib, "orc_sampler_thinned", C.c_int64, _P, _U32, _U32, C.c_int64, _D, _D, _D)
    _sig(lib, "orc_sampler_thinned_pve", C.c_int64, _P, _D)
    _sig(lib, "orc_sampler_thinned_beta", C.c_int64, _P, _D, C.c_int64)
    _sig(lib, "orc_sampler_state", None, _P, _D)
    _sig(lib, "orc_sampler_loci", C.c_int, _P, _U32)
    _sig(lib, "orc_sampler_rb", None, _P, _D, _D)
    _sig(lib, "orc_sampler_current_model", _P, _P)
    _sig(lib, "orc_sampler_new_model", _P, _P)
    _sig(lib, "orc_sampler_rng_counter", C.c_uint64, _P)
    _sig(lib, "orc_model_add_term", C.c_int, _P, C.c_uint32, C.c_int, _D)
    _sig(lib, "orc_model_remove_term", None, _P, C.c_int)
    _sig(lib, "orc_model_compute_log_likelihood", None, _P)
    _sig(lib, "orc_model_log_likelihood", C.c_double, _P)
    _sig(lib, "orc_model_size", C.c_int, _P)
    _sig(lib, "orc_model_cols", C.c_int, _P)
    _sig(lib, "orc_model_copy", None, _P, _P)
    _sig(lib, "orc_model_get", None, _P, _D, _D, _D, _D, _D, _D)
    _sig(lib, "orc_model_set_d", None, _P, C.c_int, C.c_double)
    _sig(lib, "orc_model_sample_beta_sigma2", None, _P, _P)
    _sig(lib, "orc_model_compute_pve", None, _P, _D)
    _sig(lib, "orc_model_beta", _D, _P)
    _sig(lib, "orc_model_y_hat", _D, _P)
    _sig(lib, "orc_model_model_inds", _I32, _P)
    _sig(lib, "orc_exhaustive_modelset", None, _P, C.c_int, C.c_int, _D, _D)
    _sig(lib, "orc_proposal_probs_exh", None, C.c_int, C.c_int, _U8, _D, _D, C.c_double,
         C.c_double, C.c_size_t, C.c_size_t, _D, _D)
    _sig(lib, "orc_prior_change_on_add", C.c_double, _P, _I32, C.c_int, C.c_int)
    _sig(lib, "orc_prior_change_on_rem", C.c_double, _P, _I32, C.c_int, C.c_int)
    _sig(lib, "orc_prior_change_on_swi", C.c_double, _P, _I32, C.c_int, C.c_int)
    _sig(lib, "orc_prior_log_model", C.c_double, _P, _I32)
    _sig(lib, "orc_prior_get", None, _P, _D)
    _sig(lib, "orc_sumtree_create", _P, _D, C.c_uint64)
    _sig(lib, "orc_sumtree_destroy", None, _P)
    _sig(lib, "orc_sumtree_zero", None, _P, C.c_uint64)
    _sig(lib, "orc_sumtree_unzero", None, _P, C.c_uint64)
    _sig(lib, "orc_sumtree_set_weights", None, _P, _D)
    _sig(lib, "orc_sumtree_total", C.c_double, _P)
    _sig(lib, "orc_sumtree_pick", C.c_uint64, _P, C.c_double)
    _sig(lib, "orc_sumtree_zeroed", C.c_int, _P, C.c_uint64)
    _sig(lib, "orc_synth_genotypes", None, C.c_uint32, C.c_uint32, C.c_uint64, _D, _D, _U8)
    _sig(lib, "orc_geometric_dist_cdf", None, C.c_int, C.c_double, _D)
    _sig(lib, "orc_sample_discrete_naive", C.c_int, _D, C.c_int, C.c_double)
    _sig(lib, "orc_ess_geyer", C.c_double, _D, C.c_int64)
    _sig(lib, "orc_rng_init", None, C.POINTER(Rng), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32)
    _sig(lib, "orc_rng_mvnormal_invchol", None, C.POINTER(Rng), _D, _D, C.c_int, C.c_int, C.c_double, _D)
    _sig(lib, "orc_model_draw_beta_sigma2", C.c_int64, _P, _P, C.c_int64, _D, _D)
    _sig(lib, "orc_sampler_refresh_draws", C.c_int64, _P, C.c_int64, _D, _D)
    _sig(lib, "orc_adapt_jd_mb", C.c_double, C.c_double, _D, _D, _D, C.c_int)
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def ptr(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def u8(a):
    return ptr(a, C.c_uint8)


# ------------------------------------------------------------------ helpers
def rng_fill(seed, chain, c2, c3, kind, a, b, cnt):
    out = np.empty(cnt, dtype=np.float64)
    lib().orc_rng_fill(seed, chain, c2, c3, kind, a, b, cnt, ptr(out))
    return out


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(ptr(c, C.c_uint32), ptr(k, C.c_uint32), ptr(o, C.c_uint32))
    return o


def decode(body, n, snp, coding):
    stride = (n + 3) // 4
    out = np.empty(n, dtype=np.float64)
    b = np.ascontiguousarray(body, dtype=np.uint8)
    lib().orc_get_genotypes(u8(b), stride, n, coding, snp, ptr(out))
    return out


def read_fam_pheno(path):
    y = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                y.append(float(parts[5]))
    return np.array(y, dtype=np.float64)


def read_bed(path, n, p):
    raw = open(path, "rb").read()
    assert raw[:3] == b"\x6c\x1b\x01"
    stride = (n + 3) // 4
    return np.frombuffer(raw[3:3 + stride * p], dtype=np.uint8).copy()


class Data:
    def __init__(self, n, p, body, y, e_cov=None, recode=False):
        self.n, self.p = n, p
        self.body = np.ascontiguousarray(body, dtype=np.uint8)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        m_e = 0 if e_cov is None else e_cov.shape[1]
        self._e = None if e_cov is None else np.asfortranarray(e_cov, dtype=np.float64)
        self.h = lib().orc_data_create(n, p, m_e, u8(self.body), ptr(self.y),
                                       None if self._e is None else ptr(self._e),
                                       1 if recode else 0)

    def stats(self):
        o = np.zeros(4)
        lib().orc_data_stats(self.h, ptr(o))
        return dict(yy=o[0], var_y=o[1], var_x=o[2], mean_x=o[3])

    def body_now(self):
        stride = (self.n + 3) // 4
        b = lib().orc_data_body(self.h)
        return np.ctypeslib.as_array(b, shape=(stride * self.p,)).copy()

    def missing(self, snp):
        locs = (C.c_size_t * 64)()
        pc = np.zeros(3)
        c = lib().orc_data_missing(self.h, snp, locs, 64, ptr(pc))
        return [locs[i] for i in range(c)], pc

    def __del__(self):
        try:
            lib().orc_data_destroy(self.h)
        except Exception:
            pass


class DataModel:
    def __init__(self, data, types):
        self.data = data
        tv = np.array(sorted(EF[t] for t in types), dtype=np.int32)
        self.types = tv
        self.h = lib().orc_datamodel_create(data.h, ptr(tv, C.c_int32), len(tv))

    def get(self, coding, snp):
        out = np.empty(self.data.n)
        lib().orc_datamodel_get(self.h, coding, snp, ptr(out))
        return out

    def offset(self):
        ot = (C.c_int * 4)()
        off = lib().orc_datamodel_offset(self.h, ot)
        return off, list(ot)

    def self_moments(self):
        off, _ = self.offset()
        out = np.empty(self.data.p * off)
        lib().orc_self_moments(self.h, ptr(out))
        return out.reshape(self.data.p, off)

    def __del__(self):
        try:
            lib().orc_datamodel_destroy(self.h)
        except Exception:
            pass


class Model:
    """A view of one of a sampler's models (cur or new)."""

    def __init__(self, sampler, h):
        self.s, self.h = sampler, h

    def add_term(self, ind, t_ind, d2):
        d = np.asarray(d2, dtype=np.float64)
        return lib().orc_model_add_term(self.h, ind, t_ind, ptr(d))

    def remove_term(self, model_ind):
        lib().orc_model_remove_term(self.h, model_ind)

    def compute_log_likelihood(self):
        lib().orc_model_compute_log_likelihood(self.h)

    def log_likelihood(self):
        return lib().orc_model_log_likelihood(self.h)

    def size(self):
        return lib().orc_model_size(self.h)

    def cols(self):
        return lib().orc_model_cols(self.h)

    def set_d(self, x_ind, val):
        lib().orc_model_set_d(self.h, x_ind, val)

    def get(self):
        k = self.cols()
        xx = np.zeros(k * k); l = np.zeros(k * k)
        xy = np.zeros(k); v = np.zeros(k); d = np.zeros(k); sc = np.zeros(6)
        lib().orc_model_get(self.h, ptr(xx), ptr(l), ptr(xy), ptr(v), ptr(d), ptr(sc))
        return dict(xx=xx.reshape(k, k).T, l=l.reshape(k, k).T, xy=xy, v=v, d=d,
                    loglik=sc[0], syx=sc[1], ldq=sc[2], ldl=sc[3], sigma2=sc[4])

    def copy_from(self, other):
        lib().orc_model_copy(self.h, other.h)


class Sampler:
    def __init__(self, cfg, data):
        self.opts = make_opts(cfg)
        self.data = data
        self.h = lib().orc_sampler_create(C.byref(self.opts), data.h)
        if not self.h:
            raise ValueError("invalid sampler configuration")

    def init_flat(self):
        lib().orc_sampler_init_p_proposal_flat(self.h)

    def setup(self):
        return lib().orc_sampler_setup(self.h)

    def marginal_z(self):
        z = np.zeros(self.data.p, np.float32)
        lib().orc_sampler_marginal_z(self.h, ptr(z, C.c_float))
        return z

    def swap_labels(self, H, cap=1 << 20):
        j = np.zeros(cap, np.uint32); o = np.zeros(cap, np.int32); lab = np.zeros(cap, np.float32)
        k = lib().orc_sampler_swap_labels(self.h, H, ptr(j, C.c_uint32), ptr(o, C.c_int32),
                                          ptr(lab, C.c_float), cap)
        k = min(k, cap)
        return j[:k], o[:k], lab[:k]

    def set_move_mix(self, p):
        a = np.ascontiguousarray(p, dtype=np.float64)
        return lib().orc_sampler_set_move_mix(self.h, ptr(a))

    def set_learned(self, w, H, p_learned):
        """Frozen learned-swap weights w (p x 2H float32), or None."""
        if w is None:
            return lib().orc_sampler_set_learned(self.h, None, H, p_learned)
        self._lw = np.ascontiguousarray(w, dtype=np.float32)
        assert self._lw.shape == (self.data.p, 2 * H)
        return lib().orc_sampler_set_learned(self.h, ptr(self._lw, C.c_float), H, p_learned)

    def iterate(self, n):
        rc = lib().orc_sampler_iterate(self.h, n)
        if rc:
            raise RuntimeError("sampler error %d" % rc)

    def sample(self, n):
        rc = lib().orc_sampler_sample(self.h, n)
        if rc:
            raise RuntimeError("sampler error %d" % rc)

    def trace(self):
        st = self.state()
        n = int(st["n_iter"])
        mv = np.zeros(n, np.uint8); jd = np.zeros(n, np.uint8); ms = np.zeros(n, np.uint8)
        sz = np.zeros(n, np.uint32); ll = np.zeros(n)
        lib().orc_sampler_trace(self.h, u8(mv), u8(jd), u8(ms), ptr(sz, C.c_uint32), ptr(ll))
        return dict(move=mv, jd=jd, ms=ms, size=sz, loglik=ll)

    def thinned(self):
        st = self.state()
        nl = int(st["loci_len"])
        n = int(st["n_iter"])
        sizes = np.zeros(n + 1, np.uint32)
        loci = np.zeros(max(nl, 1), np.uint32)
        ll = np.zeros(n + 1); s2 = np.zeros(n + 1); al = np.zeros(n + 1)
        cnt = lib().orc_sampler_thinned(self.h, ptr(sizes, C.c_uint32), ptr(loci, C.c_uint32),
                                        nl, ptr(ll), ptr(s2), ptr(al))
        pve = np.zeros(3 * (n + 1))
        lib().orc_sampler_thinned_pve(self.h, ptr(pve))
        nb = lib().orc_sampler_thinned_beta(self.h, None, 0)
        beta = np.zeros(max(nb, 1))
        lib().orc_sampler_thinned_beta(self.h, ptr(beta), nb)
        return dict(sizes=sizes[:cnt], loci=loci[:nl], loglik=ll[:cnt], sigma2=s2[:cnt],
                    alpha=al[:cnt], pve=pve[:3 * cnt].reshape(cnt, 3), beta=beta[:nb])

    def state(self):
        o = np.zeros(8)
        lib().orc_sampler_state(self.h, ptr(o))
        return dict(loglik=o[0], size=int(o[1]), sigma2=o[2], alpha=o[3], p_move_size=o[4],
                    n_accepted=o[5], n_iter=o[6], loci_len=o[7])

    def loci(self):
        buf = np.zeros(4096, np.uint32)
        k = lib().orc_sampler_loci(self.h, ptr(buf, C.c_uint32))
        return buf[:k].copy()

    def p_rao(self):
        return np.ctypeslib.as_array(lib().orc_sampler_p_rao(self.h), shape=(self.data.p,)).copy()

    def p_proposal(self):
        return np.ctypeslib.as_array(lib().orc_sampler_p_proposal(self.h),
                                     shape=(self.data.p,)).copy()

    def set_p_proposal(self, p):
        """Replace the add/remove proposal's source (q_add / q_rem and the
        trees follow); 0 on success."""
        self._pp = np.ascontiguousarray(p, dtype=np.float64)
        assert self._pp.shape == (self.data.p,)
        return lib().orc_sampler_set_p_proposal(self.h, ptr(self._pp, C.c_double))

    def refresh_draws(self, count):
        """count independent tau2/alpha refreshes of the current model from
        its present state (src/prior.cpp:30-141): alpha (count,) and the
        terms' inv_tau2_alpha2 (count, k)."""
        k = int(self.current().cols()) - 1
        al = np.zeros(count)
        dv = np.zeros(count * max(k, 1))
        k2 = lib().orc_sampler_refresh_draws(self.h, count, ptr(al), ptr(dv))
        return al, dv[:count * k2].reshape(count, k2)

    def beta_sigma2_draws(self, count):
        """count draws of (sigma2, beta) on the current model (src/model.hpp:
        314-342) from the chain's stream."""
        k = int(self.current().cols())
        s2 = np.zeros(count)
        b = np.zeros(count * k)
        m = lib().orc_sampler_current_model(self.h)
        k2 = lib().orc_model_draw_beta_sigma2(self.h, m, count, ptr(s2), ptr(b))
        return s2, b[:count * k2].reshape(count, k2)

    def rb(self, y_hat):
        out = np.zeros(self.data.p)
        yh = np.ascontiguousarray(y_hat, dtype=np.float64)
        lib().orc_sampler_rb(self.h, ptr(yh), ptr(out))
        return out

    def current(self):
        return Model(self, lib().orc_sampler_current_model(self.h))

    def new(self):
        return Model(self, lib().orc_sampler_new_model(self.h))

    def prior(self):
        o = np.zeros(8)
        lib().orc_prior_get(self.h, ptr(o))
        return dict(g_a=o[0], g_b=o[1], alpha=o[2], inv_tau2_alpha2_A=o[3],
                    inv_tau2_alpha2_H=o[4], nus2_plus_yy=o[5], n_plus_nu=o[6])

    def prior_add(self, Ns, n_old, t):
        a = np.asarray(Ns, dtype=np.int32)
        return lib().orc_prior_change_on_add(self.h, ptr(a, C.c_int32), n_old, t)

    def prior_rem(self, Ns, n_old, t):
        a = np.asarray(Ns, dtype=np.int32)
        return lib().orc_prior_change_on_rem(self.h, ptr(a, C.c_int32), n_old, t)

    def prior_swi(self, Ns, ta, tr):
        a = np.asarray(Ns, dtype=np.int32)
        return lib().orc_prior_change_on_swi(self.h, ptr(a, C.c_int32), ta, tr)

    def prior_log_model(self, Ns):
        a = np.asarray(Ns, dtype=np.int32)
        return lib().orc_prior_log_model(self.h, ptr(a, C.c_int32))

    def rng_counter(self):
        return lib().orc_sampler_rng_counter(self.h)

    def __del__(self):
        try:
            lib().orc_sampler_destroy(self.h)
        except Exception:
            pass


MLP_IN, MLP_HID = 16, 256
MLP_OFF = dict(W1=0, b1=MLP_IN * MLP_HID)
MLP_OFF["W2"] = MLP_OFF["b1"] + MLP_HID
MLP_OFF["b2"] = MLP_OFF["W2"] + MLP_HID * MLP_HID
MLP_OFF["w3"] = MLP_OFF["b2"] + MLP_HID
MLP_OFF["b3"] = MLP_OFF["w3"] + MLP_HID
MLP_NPAR = MLP_OFF["b3"] + 1


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mlp_forward(params, X):
    p, x = _f32(params), _f32(X)
    out = np.zeros(x.shape[0], np.float32)
    lib().orc_mlp_forward(ptr(p, C.c_float), ptr(x, C.c_float), x.shape[0], ptr(out, C.c_float))
    return out


def mlp_gradients(params, X, labels, exact=False):
    p, x, lab = _f32(params), _f32(X), _f32(labels)
    g = np.zeros(MLP_NPAR)
    f = lib().orc_mlp_loss_exact if exact else lib().orc_mlp_gradients
    loss = f(ptr(p, C.c_float), ptr(x, C.c_float), ptr(lab, C.c_float), x.shape[0], ptr(g))
    return loss, g


def mlp_loss_exact(params, X, labels):
    p, x, lab = _f32(params), _f32(X), _f32(labels)
    return lib().orc_mlp_loss_exact(ptr(p, C.c_float), ptr(x, C.c_float), ptr(lab, C.c_float),
                                    x.shape[0], None)


def mlp_adam(params, m, v, grads, lr, t):
    """In place on float32 arrays params, m, v."""
    g = _f32(grads)
    lib().orc_mlp_adam(ptr(params, C.c_float), ptr(m, C.c_float), ptr(v, C.c_float),
                       ptr(g, C.c_float), C.c_float(lr), t)


def learn_pair_features(pre, ot, n, band, H, z, pbar, j, o):
    """Features of pair (j, j + o); pre is p x offset self moments."""
    pre = np.ascontiguousarray(pre, dtype=np.float64)
    band = np.ascontiguousarray(band, dtype=np.int32)
    z = _f32(z); pbar = _f32(pbar)
    f = np.zeros(16, np.float32)
    lib().orc_learn_pair_features(ptr(pre), pre.shape[1], ot, n, ptr(band, C.c_int32), H,
                                  ptr(z, C.c_float), ptr(pbar, C.c_float), int(j), int(o),
                                  ptr(f, C.c_float))
    return f


def learn_snp_features(pre, ot, n, band, H, z, pbar, model_inds, p_proposal, j):
    """Features of SNP j for the per-SNP proposal network in a chain's state."""
    pre = np.ascontiguousarray(pre, dtype=np.float64)
    band = np.ascontiguousarray(band, dtype=np.int32)
    z = _f32(z); pbar = _f32(pbar)
    mi = np.ascontiguousarray(model_inds, dtype=np.int32)
    pp = np.ascontiguousarray(p_proposal, dtype=np.float64)
    f = np.zeros(16, np.float32)
    lib().orc_learn_snp_features(ptr(pre), pre.shape[1], ot, n, len(z), ptr(band, C.c_int32), H,
                                 ptr(z, C.c_float), ptr(pbar, C.c_float), ptr(mi, C.c_int32), ptr(pp), int(j),
                                 ptr(f, C.c_float))
    return f


def mlp_init(seed, scale3=0.1):
    """He-initialised parameters (biases small), float32."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_NPAR, np.float32)
    o = MLP_OFF
    p[o["W1"]:o["b1"]] = rng.normal(scale=np.sqrt(2.0 / MLP_IN), size=MLP_IN * MLP_HID)
    p[o["b1"]:o["W2"]] = rng.normal(scale=0.05, size=MLP_HID)
    p[o["W2"]:o["b2"]] = rng.normal(scale=np.sqrt(2.0 / MLP_HID), size=MLP_HID * MLP_HID)
    p[o["b2"]:o["w3"]] = rng.normal(scale=0.05, size=MLP_HID)
    p[o["w3"]:o["b3"]] = rng.normal(scale=scale3, size=MLP_HID)
    p[o["b3"]] = 0.1
    return p


class SumTree:
    def __init__(self, w):
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        self.h = lib().orc_sumtree_create(ptr(self.w), len(self.w))

    def zero(self, i):
        lib().orc_sumtree_zero(self.h, i)

    def unzero(self, i):
        lib().orc_sumtree_unzero(self.h, i)

    def total(self):
        return lib().orc_sumtree_total(self.h)

    def pick(self, u):
        return lib().orc_sumtree_pick(self.h, u)

    def zeroed(self, i):
        return bool(lib().orc_sumtree_zeroed(self.h, i))

    def set_weights(self, w):
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        lib().orc_sumtree_set_weights(self.h, ptr(self.w))

    def __del__(self):
        try:
            lib().orc_sumtree_destroy(self.h)
        except Exception:
            pass


def synth_genotypes(seed, n, p, thresh, rho):
    stride = (n + 3) // 4
    body = np.zeros(stride * p, dtype=np.uint8)
    t = np.ascontiguousarray(thresh, dtype=np.float64)
    r = np.ascontiguousarray(rho, dtype=np.float64)
    lib().orc_synth_genotypes(seed, n, p, ptr(t), ptr(r), u8(body))
    return body


def ess_geyer(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_ess_geyer(ptr(a), len(a))
