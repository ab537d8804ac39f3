"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/*.h, and rejects bad configurations with a status code
and a message (never an abort)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(hm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared_functions()
    for required in ("hm_ctx_create", "hm_ctx_destroy", "hm_compute_convective_flux_and_source_on_patch",
                     "hm_compute_convective_flux_and_source_on_patches", "hm_get_num_ghosts"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from hamers_amd import LIB_PATH

    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def _cfg(**kw):
    from hamers_amd import capi

    c = capi.hm_config()
    c.dim, c.model, c.num_species, c.scheme = 2, capi.HM_MODEL_SINGLE_SPECIES, 1, capi.HM_SCHEME_WCNS6_LD_HLLC_HLL
    c.constant_p, c.constant_q, c.constant_C, c.constant_alpha_tau = 2, 4, 1e9, 35.0
    c.species_gamma[0] = 1.4
    c.species_R[0] = 1.0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("kw,status", [
    ({"dim": 4}, 1),
    ({"scheme": 7}, 2),
    ({"constant_p": 0}, 1),
])
def test_bad_configuration_is_rejected_with_message(kw, status):
    from hamers_amd import capi

    ctx = ctypes.c_void_p()
    st = capi.lib.hm_ctx_create(ctypes.byref(_cfg(**kw)), ctypes.byref(ctx))
    assert st == status
    msg = capi.lib.hm_last_error(ctx)
    assert msg and len(msg) > 0
    capi.lib.hm_ctx_destroy(ctx)


def test_null_arguments_are_rejected():
    from hamers_amd import capi

    assert capi.lib.hm_ctx_create(None, None) == capi.HM_ERR_INVALID_ARGUMENT
    assert capi.lib.hm_compute_convective_flux_and_source_on_patch(None, None, 0.0, 1.0, 0, None) == \
        capi.HM_ERR_INVALID_ARGUMENT
    g = ctypes.c_int()
    assert capi.lib.hm_get_num_ghosts(None, ctypes.byref(g)) == capi.HM_ERR_INVALID_ARGUMENT


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_reported_not_aborted():
    from hamers_amd import capi

    ctx = ctypes.c_void_p()
    st = capi.lib.hm_ctx_create(ctypes.byref(_cfg()), ctypes.byref(ctx))
    assert st == capi.HM_ERR_DEVICE
    assert b"device" in capi.lib.hm_last_error(ctx)
    capi.lib.hm_ctx_destroy(ctx)


def test_reconstructor_manager_rejects_unknown_key():
    from hamers_amd import ConvectiveFluxReconstructorManager

    with pytest.raises(ValueError):
        ConvectiveFluxReconstructorManager("x", 2, "SINGLE_SPECIES", "WCNS5_Z_HLLC_HLL")
