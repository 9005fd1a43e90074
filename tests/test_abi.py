"""The C-ABI library loads and exports every entry point include/nudim.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "nudim.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nudim_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["nudim_create", "nudim_iterate", "nudim_nfft_trafo", "nudim_nfft_adjoint", "nudim_cgnr",
              "nudim_tv", "nudim_diffusion2d", "nudim_constraint", "nudim_zfft"]:
        assert s in syms


def test_library_exports_every_declared_symbol(hip):
    L = ctypes.CDLL(hip.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header(hip):
    assert sorted(hip.EXPORTS) == declared_symbols()


def test_errors_are_reported_not_raised_in_c(hip):
    # invalid parameters fail with the reference's message, no device work needed
    p = hip.make_params(64, 64, 64, -60, 60, 3, nhio=-1, ner=20)
    with pytest.raises(hip.NudimError, match="valid nhio"):
        hip.Reconstruction(p)
    p = hip.make_params(64, 64, 64, -60, 60, 3, padangle=2.0)
    with pytest.raises(hip.NudimError, match="padangle"):
        hip.Reconstruction(p)
    p = hip.make_params(64, 64, 64, -60, 60, 3, height=10)
    with pytest.raises(hip.NudimError, match="height and center"):
        hip.Reconstruction(p)


def test_estimate_bytes_scales(hip):
    small = hip.estimate_bytes(hip.make_params(128, 128, 128, -60, 60, 3))
    big = hip.estimate_bytes(hip.make_params(384, 256, 384, -70, 70, 2))
    assert big > 10 * small > 0


def test_estimate_counts_kept_directions(hip, monkeypatch):
    """The admission estimate holds the projected-start solve's ten kept
    directions (complex FP32 N^2 nz each, and ten FP64 alphas per plane)
    exactly when a job keeps them: up to 4 GiB of directions (the 256^3
    workload, 3 GB), not at np = 768, nz = 512 (24 GB) unless forced."""
    def est(np_, nz, keep):
        if keep is None:
            monkeypatch.delenv("NUDIM_KEEP_P", raising=False)
        else:
            monkeypatch.setenv("NUDIM_KEEP_P", keep)
        return hip.estimate_bytes(hip.make_params(np_, nz, np_, -70, 70, 2, weight=1, precision=hip.FP32))
    for np_, nz, kept in ((384, 256, True), (768, 512, False)):
        extra = 10 * np_ * np_ * nz * 8 + 10 * nz * 8
        assert est(np_, nz, None) - est(np_, nz, "0") == (extra if kept else 0)
        assert est(np_, nz, "1") - est(np_, nz, "0") == extra



def test_noalloc_rejects_oversize_before_device_work(hip):
    """The no-allocation TV entry point checks the volume against its fixed
    scratch on the host, before touching the device (so no GPU is needed
    here and the address is never dereferenced)."""
    with pytest.raises(hip.NudimError, match="bytes of scratch; the maximum is 71303168"):
        hip.tv_height_noalloc(0x1000, (256, 256, 256), 64, 128)
    with pytest.raises(hip.NudimError, match="bytes of scratch"):
        hip.tv_height_noalloc(0x1000, (140, 128, 128), 64, 64, precision=hip.FP64_PLAIN)


def test_tv_volume_limit_refused_up_front(hip):
    """TV jobs of 2^32 bytes or more are refused by create and by the
    admission estimate (before any device work), not at their first pass."""
    p = hip.make_params(1024, 1024, 1024, -60, 60, 3, denoise=True, method=hip.TV)
    with pytest.raises(hip.NudimError, match="fewer than 2\\^32 bytes"):
        hip.estimate_bytes(p)
    with pytest.raises(hip.NudimError, match="fewer than 2\\^32 bytes"):
        hip.Reconstruction(p)
    # the same volume in FP32 (4 GiB exactly) is refused as well; one plane pair fewer fits
    p = hip.make_params(1024, 1024, 1024, -60, 60, 3, denoise=True, method=hip.TV, precision=hip.FP32)
    with pytest.raises(hip.NudimError, match="fewer than 2\\^32 bytes"):
        hip.estimate_bytes(p)
    assert hip.estimate_bytes(hip.make_params(1024, 1022, 1024, -60, 60, 3, denoise=True, method=hip.TV,
                                              precision=hip.FP32)) > 0
    # without denoising the size is not limited by the TV kernels
    assert hip.estimate_bytes(hip.make_params(1024, 1024, 1024, -60, 60, 3)) > 0


def test_batch_arguments_checked_before_device_work(hip):
    """The batch entry points refuse a volume count outside 1..4096 and null
    contexts with an error code and message (no GPU needed)."""
    p = hip.make_params(32, 32, 32, -60, 60, 5)
    for nvol in (0, -2, 4097):
        with pytest.raises(hip.NudimError, match="1 to 4096 volumes"):
            hip.Batch(p, nvol)
    # two 512^3 volumes (np 768, n 1536): 2.4e9 grid elements, over 2^31
    big = hip.make_params(768, 512, 768, -60, 60, 1, precision=hip.FP32)
    with pytest.raises(hip.NudimError, match="below 2\\^31"):
        hip.Batch(big, 2)
    L = hip.lib()
    assert L.nudim_batch_load_projections(None, 0, None) != 0
    assert b"null argument" in L.nudim_last_error()
    assert L.nudim_batch_get_volume(None, 0, None) != 0
