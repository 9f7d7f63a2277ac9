"""ctypes binding of libbmagwa_hip.so (the C ABI of include/bmagwa_hip.h)."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libbmagwa_hip.so")
EF = {"A": 0, "H": 1, "D": 2, "R": 3, "AH": 4}


def lib_path():
    return LIB


def build(jobs=8):
    """Compile every HIP source for gfx950 into bmagwa_amd/libbmagwa_hip.so."""
    subprocess.check_call(["make", "-s", "-j%d" % jobs, "-C", os.path.join(HERE, "csrc")])


class BmgError(RuntimeError):
    pass


class Opts(C.Structure):
    """bmg_opts: the INI keys of the reference (src/options.hpp:138-321)."""
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


class BatchCfg(C.Structure):
    _fields_ = [
        ("trace_every_iter", C.c_int32), ("proposal", C.c_int32), ("learn_every", C.c_int32),
        ("learn_freeze_iter", C.c_int32), ("scan_digits", C.c_int32), ("use_graph", C.c_int32),
        ("trace_capacity", C.c_uint64),
        ("learn_window", C.c_int32), ("learn_steps", C.c_int32), ("learn_batch", C.c_int32),
        ("capacity", C.c_int32), ("p_learned", C.c_double), ("learn_rate", C.c_double),
        ("max_model_size", C.c_int32), ("cfg_pad", C.c_int32),
    ]


BATCH_CFG_DEFAULT = dict(trace_every_iter=0, proposal=0, learn_every=500, learn_freeze_iter=0,
                         scan_digits=5, use_graph=0, trace_capacity=0, learn_window=20,
                         learn_steps=100, learn_batch=4096, capacity=0, p_learned=0.9,
                         learn_rate=1e-3, max_model_size=0, cfg_pad=0)

OPT_DEFAULTS = dict(max_move_size=20, max_SNP_neighborhood_size=10, p_move_size=0.2,
                    p_move_size_nbs=0.2, p_move_size_nbc=0.2, adapt_p_move_size=1,
                    p_move_size_acpt_goal=0.0, adaptation=0, flat_proposal_dist=0,
                    save_beta=0, delay_rejection=0, R2mode_sigma2=0.0, s2_sigma2=0.0,
                    mu_alpha=0.0, inv_tau2_e_val=0.0, use_individual_tau2=0,
                    recode_g_to_minor_allele_count=0, m_e=0, chain=0, sampler_type=0)


def opts_from_cfg(cfg):
    """INI-style dict -> Opts (defaults of src/options.hpp:138-221)."""
    o = Opts()
    d = dict(OPT_DEFAULTS)
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
    d.pop("verbosity", None)
    for k, v in d.items():
        setattr(o, k, v)
    return o


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_U32 = C.POINTER(C.c_uint32)
_I32 = C.POINTER(C.c_int32)
_U64 = C.POINTER(C.c_uint64)
_I64 = C.POINTER(C.c_int64)

SIGNATURES = {
    "bmg_last_error": (C.c_char_p,),
    "bmg_device_count": (C.c_int, C.POINTER(C.c_int)),
    "bmg_ctx_create": (C.c_int, C.c_int, _U8, C.c_uint32, C.c_uint64, _D, _I32, C.c_int, C.c_int, _PP),
    "bmg_ctx_create_synthetic": (C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint64, _D, _D, _I32,
                                 C.c_int, _PP),
    "bmg_ctx_set_y": (C.c_int, _P, _D),
    "bmg_ctx_set_covariates": (C.c_int, _P, _D, C.c_uint32),
    "bmg_ctx_create_subset": (C.c_int, _P, _U32, C.c_uint32, _D, _I32, C.c_int, _PP),
    "bmg_ctx_destroy": (C.c_int, _P),
    "bmg_ctx_read_body": (C.c_int, _P, _U8),
    "bmg_ctx_stats": (C.c_int, _P, _D),
    "bmg_ctx_missing_count": (C.c_int, _P, _U64),
    "bmg_decode": (C.c_int, _P, C.c_uint64, C.c_int, _D),
    "bmg_decode_many": (C.c_int, _P, _U64, C.c_uint64, C.c_int, _D),
    "bmg_self_moments_layout": (C.c_int, _P, C.POINTER(C.c_int), C.POINTER(C.c_int)),
    "bmg_self_moments": (C.c_int, _P, _D),
    "bmg_band_xtx": (C.c_int, _P, C.c_uint32, C.c_int, C.c_int, _I32),
    "bmg_band_xtx_dev": (C.c_int, _P, C.c_uint32, C.c_int, C.c_int, _P, _P),
    "bmg_scan_xtr": (C.c_int, _P, _D, C.c_int, C.c_int, C.c_int, _D),
    "bmg_scan_xtr_dev": (C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P),
    "bmg_scan_xtr_dev_noalloc": (C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P),
    "bmg_ctx_n_pad": (C.c_int, _P, _U32),
    "bmg_scan_xtr_timed": (C.c_int, _P, _P, C.c_int, C.c_int, C.c_int, _P, C.c_int, _D, _P),
    "bmg_scan_shape": (C.c_int, _P, C.c_int, C.c_int, _P),
    "bmg_scan_set_shape": (C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32),
    "bmg_band_xtx_timed": (C.c_int, _P, C.c_uint32, C.c_int, C.c_int, _P, C.c_int, _D, _P),
    "bmg_dev_alloc": (C.c_int, C.c_int, C.c_uint64, _PP),
    "bmg_dev_free": (C.c_int, C.c_int, _P),
    "bmg_memcpy_h2d": (C.c_int, C.c_int, _P, _P, C.c_uint64),
    "bmg_memcpy_d2h": (C.c_int, C.c_int, _P, _P, C.c_uint64),
    "bmg_copy_timed": (C.c_int, C.c_int, _P, _P, C.c_uint64, C.c_int, _D),
    "bmg_read_timed": (C.c_int, C.c_int, _P, C.c_uint64, _P, C.c_int, _D),
    "bmg_mlp_nparams": (C.c_int,),
    "bmg_mlp_create": (C.c_int, C.c_int, _PP),
    "bmg_mlp_destroy": (C.c_int, _P),
    "bmg_mlp_set_params": (C.c_int, _P, C.POINTER(C.c_float)),
    "bmg_mlp_get_params": (C.c_int, _P, C.POINTER(C.c_float)),
    "bmg_mlp_forward": (C.c_int, _P, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)),
    "bmg_mlp_gradients": (C.c_int, _P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                          C.POINTER(C.c_float)),
    "bmg_mlp_adam_step": (C.c_int, _P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                          C.c_float),
    "bmg_model_kmax": (C.c_int,),
    "bmg_model_create": (C.c_int, _P, _D, _PP),
    "bmg_model_destroy": (C.c_int, _P),
    "bmg_model_create_sampler": (C.c_int, _P, C.POINTER(Opts), _PP),
    "bmg_model_sample_beta_sigma2": (C.c_int, _P),
    "bmg_model_compute_pve": (C.c_int, _P, _D),
    "bmg_model_beta": (C.c_int, _P, _D, _D, _D, C.POINTER(C.c_int)),
    "bmg_model_rb": (C.c_int, _P, _D, _D),
    "bmg_model_commit": (C.c_int, _P, _P),
    "bmg_model_rollback": (C.c_int, _P, _P),
    "bmg_model_add_term": (C.c_int, _P, C.c_uint32, C.c_int, _D),
    "bmg_model_remove_term": (C.c_int, _P, C.c_int),
    "bmg_model_compute_log_likelihood": (C.c_int, _P),
    "bmg_model_log_likelihood": (C.c_int, _P, _D),
    "bmg_model_size": (C.c_int, _P, C.POINTER(C.c_int)),
    "bmg_model_copy": (C.c_int, _P, _P),
    "bmg_model_swapadj": (C.c_int, _P, C.c_int),
    "bmg_model_get": (C.c_int, _P, _D, _D, _D, _D, _D, _D),
    "bmg_model_set_d": (C.c_int, _P, C.c_int, C.c_double),
    "bmg_batch_create": (C.c_int, _PP, C.c_int, C.c_int, _I32, C.POINTER(Opts), C.POINTER(BatchCfg), _PP),
    "bmg_batch_destroy": (C.c_int, _P),
    "bmg_batch_setup": (C.c_int, _P),
    "bmg_batch_run": (C.c_int, _P, C.c_uint64),
    "bmg_batch_sync": (C.c_int, _P),
    "bmg_batch_set_learned": (C.c_int, _P, C.c_int, C.POINTER(C.c_float), C.c_int, C.c_double),
    "bmg_batch_learn_state": (C.c_int, _P, C.c_int, _D),
    "bmg_batch_learn_table": (C.c_int, _P, C.c_int, C.POINTER(C.c_float)),
    "bmg_batch_learn_samples": (C.c_int, _P, C.c_int, _U32, _I32, C.POINTER(C.c_float), C.c_int64, _I64),
    "bmg_batch_learn_params": (C.c_int, _P, C.c_int, C.POINTER(C.c_float)),
    "bmg_batch_learn_static": (C.c_int, _P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), _I32),
    "bmg_batch_learn_features": (C.c_int, _P, C.c_int, _U32, _I32, C.c_int64, C.POINTER(C.c_float)),
    "bmg_batch_learn_snp_features": (C.c_int, _P, C.c_int, C.c_int, _U32, C.c_int64, C.POINTER(C.c_float)),
    "bmg_batch_state": (C.c_int, _P, C.c_int, _D),
    "bmg_batch_loci": (C.c_int, _P, C.c_int, _U32, C.c_int, C.POINTER(C.c_int)),
    "bmg_batch_p_rao": (C.c_int, _P, C.c_int, _D),
    "bmg_batch_p_proposal": (C.c_int, _P, C.c_int, _D),
    "bmg_batch_set_p_proposal": (C.c_int, _P, C.c_int, _D),
    "bmg_batch_trace": (C.c_int, _P, C.c_int, _U8, _U8, _U8, _U32, _D, C.c_int64, _I64),
    "bmg_batch_write_outputs": (C.c_int, _P, C.c_int, C.c_char_p),
    "bmg_batch_thinned_beta": (C.c_int, _P, C.c_int, _D, C.c_int64, _I64),
    "bmg_batch_thinned": (C.c_int, _P, C.c_int, _U32, C.c_int64, _U32, C.c_int64, _D, _D, _D, _I64, _I64),
    "bmg_batch_rb": (C.c_int, _P, C.c_int, _D, _D),
    "bmg_batch_timing": (C.c_int, _P, _D),
    "bmg_batch_graph_stats": (C.c_int, _P, _D),
    "bmg_batch_profile": (C.c_int, _P, C.c_int, _D),
}

_lib = None


def lib():
    """Load the engine; raises if it was not built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise BmgError("libbmagwa_hip.so not built (run __graft_entry__.build())")
        _lib = C.CDLL(LIB)
        for name, sig in SIGNATURES.items():
            f = getattr(_lib, name)
            f.restype = sig[0]
            f.argtypes = list(sig[1:])
    return _lib


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


def check(rc):
    if rc != 0:
        msg = lib().bmg_last_error().decode()
        raise BmgError("bmg error %d: %s" % (rc, msg))


class Ctx:
    """One dataset in HBM: Data + DataModel + PrecomputedSNPCovariances."""

    def __init__(self, n, p, body=None, y=None, types=("A",), recode=False, device=0,
                 synth=None, e_cov=None):
        self.n, self.p = n, p
        tv = np.array(sorted(EF[t] for t in types), dtype=np.int32)
        self.types = tv
        h = C.c_void_p()
        if synth is not None:
            seed, thresh, rho = synth
            t = np.ascontiguousarray(thresh, dtype=np.float64)
            r = np.ascontiguousarray(rho, dtype=np.float64)
            check(lib().bmg_ctx_create_synthetic(device, seed, n, p, _p(t), _p(r), _p(tv, C.c_int32),
                                                 len(tv), C.byref(h)))
        else:
            b = np.ascontiguousarray(body, dtype=np.uint8)
            yy = np.ascontiguousarray(y, dtype=np.float64)
            check(lib().bmg_ctx_create(device, _p(b, C.c_uint8), n, p, _p(yy), _p(tv, C.c_int32),
                                       len(tv), 1 if recode else 0, C.byref(h)))
        self.h = h
        if e_cov is not None:
            self.set_covariates(e_cov)

    def set_covariates(self, e_cov):
        """Covariates n x m_e (the constant column is added by the engine)."""
        e = np.asfortranarray(np.asarray(e_cov, dtype=np.float64).reshape(self.n, -1))
        check(lib().bmg_ctx_set_covariates(self.h, e.ctypes.data_as(C.POINTER(C.c_double)), e.shape[1]))

    @classmethod
    def subset(cls, pool, individuals, y, types=("A",)):
        """The dataset of a phenotype measured on pool individuals
        `individuals` (gathered on the device) with its own y."""
        self = cls.__new__(cls)
        ind = np.ascontiguousarray(individuals, dtype=np.uint32)
        self.n, self.p = len(ind), pool.p
        self.types = np.array(sorted(EF[t] for t in types), dtype=np.int32)
        yy = np.ascontiguousarray(y, dtype=np.float64)
        assert len(yy) == self.n
        h = C.c_void_p()
        check(lib().bmg_ctx_create_subset(pool.h, _p(ind, C.c_uint32), self.n, _p(yy),
                                          _p(self.types, C.c_int32), len(self.types), C.byref(h)))
        self.h = h
        return self

    def set_y(self, y):
        yy = np.ascontiguousarray(y, dtype=np.float64)
        check(lib().bmg_ctx_set_y(self.h, _p(yy)))

    def read_body(self):
        out = np.zeros(((self.n + 3) // 4) * self.p, dtype=np.uint8)
        check(lib().bmg_ctx_read_body(self.h, _p(out, C.c_uint8)))
        return out

    def stats(self):
        o = np.zeros(4)
        check(lib().bmg_ctx_stats(self.h, _p(o)))
        return dict(yy=o[0], var_y=o[1], var_x=o[2], mean_x=o[3])

    def missing_count(self):
        v = C.c_uint64()
        check(lib().bmg_ctx_missing_count(self.h, C.byref(v)))
        return v.value

    def decode(self, snps, coding):
        s = np.ascontiguousarray(np.atleast_1d(snps), dtype=np.uint64)
        out = np.zeros((len(s), self.n))
        check(lib().bmg_decode_many(self.h, _p(s, C.c_uint64), len(s), coding, _p(out)))
        return out

    def self_moments(self):
        off = C.c_int()
        ot = (C.c_int * 4)()
        check(lib().bmg_self_moments_layout(self.h, C.byref(off), ot))
        out = np.zeros(self.p * off.value)
        check(lib().bmg_self_moments(self.h, _p(out)))
        return out.reshape(self.p, off.value)

    def band_xtx(self, W, ca=0, cb=0):
        out = np.zeros(self.p * W, dtype=np.int32)
        check(lib().bmg_band_xtx(self.h, W, ca, cb, _p(out, C.c_int32)))
        return out.reshape(self.p, W)

    def scan_shape(self, nvec=1, digits=5):
        """Autotuned scan shape for nvec * digits digit columns: dict with
        tiles_per_wave, tile_words, n_tiles, workgroups, waves (zeros before
        the first scan of that width)."""
        out = np.zeros(5, dtype=np.int32)
        check(lib().bmg_scan_shape(self.h, nvec, digits, _p(out, C.c_int32)))
        return dict(zip(("tiles_per_wave", "tile_words", "n_tiles", "workgroups", "waves"), out.tolist()))

    def scan_set_shape(self, nvec, digits, tiles_per_wave, tile_words, workgroups, waves=4):
        """Pin the scan shape for nvec * digits columns (skips the tuner)."""
        check(lib().bmg_scan_set_shape(self.h, nvec, digits, tiles_per_wave, waves, tile_words, workgroups))

    def band_timed(self, W, ca, cb, out_dev, reps=3, stream=None):
        """Mean duration (ms) of the band kernel over `reps` launches."""
        ms = C.c_double()
        check(lib().bmg_band_xtx_timed(self.h, W, ca, cb, C.c_void_p(out_dev), reps, C.byref(ms),
                                       C.c_void_p(stream) if stream else None))
        return ms.value

    def band_xtx_dev(self, W, ca, cb, out_dev, stream=None):
        """Band X^T X into a device buffer (p x W int32) on `stream`."""
        check(lib().bmg_band_xtx_dev(self.h, W, ca, cb, C.c_void_p(out_dev),
                                     C.c_void_p(stream) if stream else None))

    def scan_xtr(self, r, coding=0, digits=5):
        r = np.ascontiguousarray(np.atleast_2d(r), dtype=np.float64)
        out = np.zeros((r.shape[0], self.p))
        check(lib().bmg_scan_xtr(self.h, _p(r), r.shape[0], coding, digits, _p(out)))
        return out

    def scan_timed(self, r_dev, nvec, out_dev, coding=0, digits=5, reps=20, stream=None):
        """Mean duration (ms) of the MFMA scan kernel over `reps` launches on
        device buffers r_dev (nvec x n_pad doubles) -> out_dev (nvec x p)."""
        ms = C.c_double()
        check(lib().bmg_scan_xtr_timed(self.h, C.c_void_p(r_dev), nvec, coding, digits,
                                       C.c_void_p(out_dev), reps, C.byref(ms),
                                       C.c_void_p(stream) if stream else None))
        return ms.value

    def scan_xtr_dev(self, r_dev, nvec, out_dev, coding=0, digits=5, stream=None, noalloc=False):
        """X^T R on device buffers r_dev (nvec x n_pad doubles) -> out_dev
        (nvec x p doubles); noalloc: the allocation-free entry point."""
        f = lib().bmg_scan_xtr_dev_noalloc if noalloc else lib().bmg_scan_xtr_dev
        check(f(self.h, C.c_void_p(r_dev), nvec, coding, digits, C.c_void_p(out_dev),
                C.c_void_p(stream) if stream else None))

    def n_pad(self):
        v = C.c_uint32()
        check(lib().bmg_ctx_n_pad(self.h, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            lib().bmg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DevBuf:
    """A device allocation made through the engine's own runtime."""

    def __init__(self, nbytes, device=0):
        self.device, self.nbytes = device, int(nbytes)
        h = C.c_void_p()
        check(lib().bmg_dev_alloc(device, self.nbytes, C.byref(h)))
        self.ptr = h.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than the buffer")
        check(lib().bmg_memcpy_h2d(self.device, C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                   a.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("download larger than the buffer")
        check(lib().bmg_memcpy_d2h(self.device, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                   out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().bmg_dev_free(self.device, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Mlp:
    """The learned-proposal network's device kernels, driven from the host."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().bmg_mlp_create(device, C.byref(h)))
        self.h = h
        self.npar = lib().bmg_mlp_nparams()

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.shape == (self.npar,)
        check(lib().bmg_mlp_set_params(self.h, _p(a, C.c_float)))

    def get_params(self):
        a = np.zeros(self.npar, np.float32)
        check(lib().bmg_mlp_get_params(self.h, _p(a, C.c_float)))
        return a

    def forward(self, X):
        x = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(x.shape[0], np.float32)
        check(lib().bmg_mlp_forward(self.h, _p(x, C.c_float), x.shape[0], _p(out, C.c_float)))
        return out

    def gradients(self, X, labels):
        x = np.ascontiguousarray(X, dtype=np.float32)
        lab = np.ascontiguousarray(labels, dtype=np.float32)
        g = np.zeros(self.npar, np.float32)
        check(lib().bmg_mlp_gradients(self.h, _p(x, C.c_float), _p(lab, C.c_float), x.shape[0],
                                      _p(g, C.c_float)))
        return g

    def adam_step(self, X, labels, lr):
        x = np.ascontiguousarray(X, dtype=np.float32)
        lab = np.ascontiguousarray(labels, dtype=np.float32)
        check(lib().bmg_mlp_adam_step(self.h, _p(x, C.c_float), _p(lab, C.c_float), x.shape[0],
                                      C.c_float(lr)))

    def close(self):
        if getattr(self, "h", None):
            lib().bmg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_timed(src, dst, nbytes, reps=10, device=0):
    """Mean duration (ms) of a plain device copy of nbytes."""
    ms = C.c_double()
    check(lib().bmg_copy_timed(device, C.c_void_p(src.ptr), C.c_void_p(dst.ptr), nbytes, reps,
                               C.byref(ms)))
    return ms.value


def read_timed(src, nbytes, sink, reps=5, device=0):
    """Best mean duration (ms) of a streaming read of nbytes over a few
    grid sizes; sink: a device buffer of at least 4 bytes."""
    ms = C.c_double()
    check(lib().bmg_read_timed(device, C.c_void_p(src.ptr), nbytes, C.c_void_p(sink.ptr), reps,
                               C.byref(ms)))
    return ms.value


class Model:
    """Drop-in Model (src/model.hpp:46-582) on the device."""

    def __init__(self, ctx, n_plus_nu=None, nus2_plus_yy=None, inv_tau2_e0=0.0, cfg=None):
        """Model(data, data_model, prior): either the solved prior terms, or
        the options of a Sampler (cfg: INI-style dict), whose prior and
        random stream the model then carries."""
        h = C.c_void_p()
        if cfg is not None:
            self._opts = opts_from_cfg(cfg)
            check(lib().bmg_model_create_sampler(ctx.h, C.byref(self._opts), C.byref(h)))
        else:
            pr = np.array([n_plus_nu, -0.5 * n_plus_nu, nus2_plus_yy, inv_tau2_e0])
            check(lib().bmg_model_create(ctx.h, _p(pr), C.byref(h)))
        self.h = h
        self.ctx = ctx

    def sample_beta_sigma2(self):
        check(lib().bmg_model_sample_beta_sigma2(self.h))

    def compute_pve(self):
        out = np.zeros(3)
        check(lib().bmg_model_compute_pve(self.h, _p(out)))
        return out

    def beta(self):
        """beta, sigma2 and y_hat of the last draws / PVE."""
        k = C.c_int()
        check(lib().bmg_model_beta(self.h, None, None, None, C.byref(k)))
        b = np.zeros(max(k.value, 1))
        s2 = C.c_double()
        yh = np.zeros(self.ctx.n)
        check(lib().bmg_model_beta(self.h, _p(b), C.byref(s2), _p(yh), C.byref(k)))
        return b[:k.value], s2.value, yh

    def rb(self, y_hat):
        out = np.zeros(self.ctx.p)
        yh = np.ascontiguousarray(y_hat, dtype=np.float64)
        check(lib().bmg_model_rb(self.h, _p(yh), _p(out)))
        return out

    def commit_from(self, proposed):
        check(lib().bmg_model_commit(self.h, proposed.h))

    def rollback_from(self, current):
        check(lib().bmg_model_rollback(self.h, current.h))

    def add_term(self, ind, t_ind, d):
        dd = np.zeros(2)
        dd[:len(np.atleast_1d(d))] = np.atleast_1d(d)
        check(lib().bmg_model_add_term(self.h, ind, t_ind, _p(dd)))

    def remove_term(self, model_ind):
        check(lib().bmg_model_remove_term(self.h, model_ind))

    def compute_log_likelihood(self):
        check(lib().bmg_model_compute_log_likelihood(self.h))

    def log_likelihood(self):
        v = C.c_double()
        check(lib().bmg_model_log_likelihood(self.h, C.byref(v)))
        return v.value

    def size(self):
        v = C.c_int()
        check(lib().bmg_model_size(self.h, C.byref(v)))
        return v.value

    def swapadj(self, col):
        check(lib().bmg_model_swapadj(self.h, col))

    def set_d(self, x_ind, val):
        check(lib().bmg_model_set_d(self.h, x_ind, val))

    def copy_from(self, other):
        check(lib().bmg_model_copy(self.h, other.h))

    def get(self):
        sc = np.zeros(6)
        check(lib().bmg_model_get(self.h, None, None, None, None, None, _p(sc)))
        k = int(sc[5])
        xx = np.zeros(k * k); l = np.zeros(k * k)
        xy = np.zeros(k); v = np.zeros(k); d = np.zeros(k)
        check(lib().bmg_model_get(self.h, _p(xx), _p(l), _p(xy), _p(v), _p(d), _p(sc)))
        return dict(xx=xx.reshape(k, k).T, l=l.reshape(k, k).T, xy=xy, v=v, d=d, loglik=sc[0],
                    syx=sc[1], ldq=sc[2], ldl=sc[3], sigma2=sc[4])

    def close(self):
        if getattr(self, "h", None):
            lib().bmg_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A set of PMV chains advanced on the device (Sampler x chains)."""

    def __init__(self, ctxs, chain_ctx, cfgs, batch_cfg=None):
        self.ctxs = list(ctxs)
        n = len(cfgs)
        arr = (C.c_void_p * len(self.ctxs))(*[c.h.value for c in self.ctxs])
        self._opts = (Opts * n)(*[opts_from_cfg(c) for c in cfgs])
        cc = np.ascontiguousarray(chain_ctx, dtype=np.int32)
        bc = dict(BATCH_CFG_DEFAULT)
        bc.update(batch_cfg or {})
        self._bcfg = BatchCfg(**bc)
        h = C.c_void_p()
        check(lib().bmg_batch_create(C.cast(arr, _PP), len(self.ctxs), n, _p(cc, C.c_int32),
                                     self._opts, C.byref(self._bcfg), C.byref(h)))
        self.h = h
        self.n_chains = n
        self.p = [self.ctxs[i].p for i in chain_ctx]

    def setup(self):
        check(lib().bmg_batch_setup(self.h))

    def run(self, n_iter):
        check(lib().bmg_batch_run(self.h, n_iter))

    def sync(self):
        check(lib().bmg_batch_sync(self.h))

    def learn_state(self, ds=0):
        o = np.zeros(4)
        check(lib().bmg_batch_learn_state(self.h, ds, _p(o)))
        return dict(rounds=int(o[0]), frozen=bool(o[1]), labels=int(o[2]), H=int(o[3]))

    def learn_table(self, ds=0):
        st = self.learn_state(ds)
        w = np.zeros((self.ctxs[ds].p, 2 * st["H"]), np.float32)
        check(lib().bmg_batch_learn_table(self.h, ds, _p(w, C.c_float)))
        return w

    def learn_samples(self, ds=0, cap=1 << 19):
        j = np.zeros(cap, np.uint32); o = np.zeros(cap, np.int32); lab = np.zeros(cap, np.float32)
        cnt = C.c_int64()
        check(lib().bmg_batch_learn_samples(self.h, ds, _p(j, C.c_uint32), _p(o, C.c_int32),
                                            _p(lab, C.c_float), cap, C.byref(cnt)))
        k = min(cnt.value, cap)
        return j[:k], o[:k], lab[:k], cnt.value

    def learn_params(self, ds=0):
        a = np.zeros(lib().bmg_mlp_nparams(), np.float32)
        check(lib().bmg_batch_learn_params(self.h, ds, _p(a, C.c_float)))
        return a

    def learn_static(self, ds=0):
        st = self.learn_state(ds)
        p = self.ctxs[ds].p
        z = np.zeros(p, np.float32); pb = np.zeros(p, np.float32)
        band = np.zeros((p, st["H"] + 1), np.int32)
        check(lib().bmg_batch_learn_static(self.h, ds, _p(z, C.c_float), _p(pb, C.c_float),
                                           _p(band, C.c_int32)))
        return z, pb, band

    def learn_features(self, ds, j, off):
        jj = np.ascontiguousarray(j, dtype=np.uint32)
        oo = np.ascontiguousarray(off, dtype=np.int32)
        out = np.zeros((len(jj), 16), np.float32)
        check(lib().bmg_batch_learn_features(self.h, ds, _p(jj, C.c_uint32), _p(oo, C.c_int32),
                                             len(jj), _p(out, C.c_float)))
        return out

    def learn_snp_features(self, ds, chain, j):
        """Per-SNP proposal features of SNPs j in chain's present state."""
        jj = np.ascontiguousarray(j, dtype=np.uint32)
        out = np.zeros((len(jj), 16), np.float32)
        check(lib().bmg_batch_learn_snp_features(self.h, ds, chain, _p(jj, C.c_uint32), len(jj),
                                                 _p(out, C.c_float)))
        return out

    def set_learned(self, ds, w, H, p_learned):
        """Frozen learned-swap weights (p x 2H float32) for dataset ds, or None."""
        if w is None:
            check(lib().bmg_batch_set_learned(self.h, ds, None, H, p_learned))
            return
        a = np.ascontiguousarray(w, dtype=np.float32)
        if a.shape != (self.ctxs[ds].p, 2 * H):
            raise ValueError("weights must be p x 2H")
        check(lib().bmg_batch_set_learned(self.h, ds, _p(a, C.c_float), H, p_learned))

    def state(self, c):
        o = np.zeros(8)
        check(lib().bmg_batch_state(self.h, c, _p(o)))
        return dict(loglik=o[0], size=int(o[1]), sigma2=o[2], alpha=o[3], p_move_size=o[4],
                    n_accepted=o[5], n_iter=o[6], loci_len=o[7])

    def loci(self, c):
        buf = np.zeros(4096, np.uint32)
        k = C.c_int()
        check(lib().bmg_batch_loci(self.h, c, _p(buf, C.c_uint32), 4096, C.byref(k)))
        return buf[:k.value].copy()

    def p_rao(self, c):
        out = np.zeros(self.p[c])
        check(lib().bmg_batch_p_rao(self.h, c, _p(out)))
        return out

    def p_proposal(self, c):
        out = np.zeros(self.p[c])
        check(lib().bmg_batch_p_proposal(self.h, c, _p(out)))
        return out

    def set_p_proposal(self, c, p):
        """Replace chain c's add/remove proposal source (q_add, q_rem and the
        proposal trees follow)."""
        a = np.ascontiguousarray(p, dtype=np.float64)
        if a.shape != (self.p[c],):
            raise ValueError("proposal table must have one value per SNP")
        check(lib().bmg_batch_set_p_proposal(self.h, c, _p(a)))

    def trace(self, c, cap=10 ** 7):
        st = self.state(c)
        n = min(int(st["n_iter"]), cap)
        mv = np.zeros(n, np.uint8); jd = np.zeros(n, np.uint8); ms = np.zeros(n, np.uint8)
        sz = np.zeros(n, np.uint32); ll = np.zeros(n)
        ln = C.c_int64()
        check(lib().bmg_batch_trace(self.h, c, _p(mv, C.c_uint8), _p(jd, C.c_uint8),
                                    _p(ms, C.c_uint8), _p(sz, C.c_uint32), _p(ll), n, C.byref(ln)))
        k = ln.value
        return dict(move=mv[:k], jd=jd[:k], ms=ms[:k], size=sz[:k], loglik=ll[:k])

    def thinned(self, c):
        ns = C.c_int64(); nl = C.c_int64()
        check(lib().bmg_batch_thinned(self.h, c, None, 0, None, 0, None, None, None,
                                      C.byref(ns), C.byref(nl)))
        sizes = np.zeros(ns.value, np.uint32); loci = np.zeros(max(nl.value, 1), np.uint32)
        ll = np.zeros(ns.value); s2 = np.zeros(ns.value); al = np.zeros(ns.value)
        check(lib().bmg_batch_thinned(self.h, c, _p(sizes, C.c_uint32), ns.value,
                                      _p(loci, C.c_uint32), nl.value, _p(ll), _p(s2), _p(al),
                                      C.byref(ns), C.byref(nl)))
        return dict(sizes=sizes, loci=loci[:nl.value], loglik=ll, sigma2=s2, alpha=al)

    def thinned_beta(self, c):
        """save_beta: the beta of every thinned sample, concatenated."""
        n = C.c_int64()
        check(lib().bmg_batch_thinned_beta(self.h, c, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        check(lib().bmg_batch_thinned_beta(self.h, c, _p(out), n.value, C.byref(n)))
        return out[:n.value]

    def rb(self, c, y_hat):
        yh = np.ascontiguousarray(y_hat, dtype=np.float64)
        out = np.zeros(self.p[c])
        check(lib().bmg_batch_rb(self.h, c, _p(yh), _p(out)))
        return out

    def profile(self, c):
        o = np.zeros(24)
        check(lib().bmg_batch_profile(self.h, c, _p(o)))
        keys = ["tau2", "addrem", "swap", "statechange", "learned", "beta_sigma2", "pve_thin", "total",
                "add_xy", "add_xtx", "add_trsv", "add_rest", "n_add", "n_rem", "rem", "refresh_draws",
                "rem_downdate", "rem_solve", "dr", "dr_score", "ar_picks", "ar_apply", "ar_reverse",
                "ar_decide"]
        return dict(zip(keys, o))

    def timing(self):
        o = np.zeros(4)
        check(lib().bmg_batch_timing(self.h, _p(o)))
        return dict(total_ms=o[0], chain_ms=o[1], scan_ms=o[2], n_scans=int(o[3]))

    def write_outputs(self, c, basename):
        """The reference's binary output files of chain c (basename + "_loci.dat", ...)."""
        check(lib().bmg_batch_write_outputs(self.h, c, basename.encode()))

    def graph_stats(self):
        """Steady-state graph: segments replayed, ms in replays, builds."""
        o = np.zeros(3)
        check(lib().bmg_batch_graph_stats(self.h, _p(o)))
        return dict(replays=int(o[0]), graph_ms=o[1], builds=int(o[2]))

    def close(self):
        if getattr(self, "h", None):
            lib().bmg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
