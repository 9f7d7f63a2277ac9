"""bmagwa_amd -- MI355X (gfx950) engine for BMAGWA's MCMC step over SNP inclusion.

The engine is the shared library libbmagwa_hip.so (C ABI in
include/bmagwa_hip.h).  This package binds it with ctypes; importing it does
not need a GPU, but every call that computes does, and there is no CPU
fallback: a missing library raises at load time.
"""
from ._lib import (BmgError, Ctx, Batch, Model, Mlp, DevBuf, build, copy_timed, read_timed, lib,
                   opts_from_cfg, lib_path, BATCH_CFG_DEFAULT)

__all__ = ["BmgError", "Ctx", "Batch", "Model", "Mlp", "DevBuf", "build", "copy_timed", "read_timed",
           "lib", "opts_from_cfg", "lib_path", "BATCH_CFG_DEFAULT"]
