"""CPU-side checks of the product library: it loads, exports every symbol that
include/inspire_hip.h declares, and its setup plan (host logic, keyword.rs:145-233)
reproduces the reference's published shapes and memory figures."""
import ctypes

import pytest

import inspire_amd as IA


def test_library_exports_every_declared_symbol():
    lib = IA.lib()
    names = IA.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_header_declares_reference_operator_surface():
    # the extern "C" names bound by src/pir/gpu/{gemv,packing_online,encode}.rs
    ref = ["gpu_gemv_upload_db", "gpu_gemv_set_device_db", "gpu_gemv_free_db", "gpu_gemv_query",
           "gpu_rotation_upload_tables", "gpu_rotation_expand_keys", "gpu_rotation_free",
           "gpu_packing_online_ensure_key_buffers", "gpu_packing_online_get_d_y_all",
           "gpu_packing_online_get_d_y_bar_all", "gpu_packing_online_upload_z_body",
           "gpu_packing_online_upload_precomp", "gpu_packing_online_adopt_precomp",
           "gpu_packing_online_upload_keys", "gpu_packing_online_compute",
           "gpu_packing_online_swap_precomp", "gpu_packing_online_free", "gpu_init_memory_pool",
           "gpu_available_memory_bytes"]
    assert set(ref) <= set(IA.exported_symbols())


def test_plan_reproduces_perf_md_5m_estimate():
    # perf.md:36-55: 5M ML-KEM-768 keys (entry 1216 B), buckets 8388608, dim0 65536
    p = IA.plan_shape(5_000_000, 32, 1184, buckets=8388608, dim0=65536)
    assert (p["db_rows"], p["db_cols"], p["instances"]) == (65536, 131072, 64)
    assert p["peak_bytes"] == 32_205_963_264


def test_plan_reproduces_perf_md_default_dim0_rejection_figure():
    # perf.md:69: the default 5M shape (default buckets 16,777,216, dim0 2048)
    # would need 669,207,494,656 bytes
    p = IA.plan_shape(5_000_000, 32, 1184)
    assert p["num_items"] == 16_777_216 and p["dim0"] == 2048
    assert p["peak_bytes"] == 669_207_494_656


def test_plan_10k_shape():
    p = IA.plan_shape(10_000, 32, 1184)
    assert p["num_items"] == 32768
    assert (p["db_rows"], p["instances"], p["interpolate_degree"], p["c"]) == (2048, 8, 8, 1)


def test_plan_1m_shape():
    p = IA.plan_shape(1_000_000, 32, 1184)
    assert p["num_items"] == 2_097_152
    assert (p["db_rows"], p["db_cols"], p["instances"]) == (2048, 720_896, 352)
    assert (p["interpolate_degree"], p["c"]) == (32, 11)
    assert p["db_bytes"] == 2_952_790_016
    assert p["bold_t_bytes"] == 2 * 352 * 1023 * 3 * 2048 * 8


def test_plan_rejects_buckets_below_record_count():
    # keyword.rs:171-180: the 1M-key / 262144-bucket configuration is rejected
    with pytest.raises(IA.InspireError) as e:
        IA.plan_shape(1_000_000, 32, 1184, buckets=262144)
    assert "bucket count 262144 must be >= record count 1000000" in str(e.value)


def test_plan_other_errors():
    with pytest.raises(IA.InspireError, match="at least one record"):
        IA.plan_shape(0, 32, 1184)
    with pytest.raises(IA.InspireError, match="multiple of 16 bits"):
        IA.plan_shape(100, 32, 1185)
    with pytest.raises(IA.InspireError, match="too small for the PIR scheme"):
        IA.plan_shape(600, 32, 1184, buckets=1024)


@pytest.mark.parametrize("kem,pk", [("512", 800), ("1024", 1568)])
def test_plan_mixed_db_shapes(kem, pk):
    if kem == "512":
        p = IA.plan_shape(4_000_000, 32, pk)
        assert (p["num_items"], p["instances"], p["c"]) == (8_388_608, 1024, 32)
    else:
        p = IA.plan_shape(250_000, 32, pk)
        assert (p["num_items"], p["instances"], p["c"]) == (524_288, 128, 4)


def test_server_rejects_bad_shapes_before_touching_a_device():
    # The shape checks come before any HIP call, so they run without a GPU.
    with pytest.raises(IA.InspireError, match="multiple of 2048"):
        IA.Server(3000, 352, 32)
    with pytest.raises(IA.InspireError, match="multiple of interpolate_degree"):
        IA.Server(2048, 350, 32)
    # 65536 outputs bound the packing's 32-bit fragment offsets
    with pytest.raises(IA.InspireError, match="at most 65536"):
        IA.Server(2048, 65536 + 32, 32)
    with pytest.raises(IA.InspireError, match="max_queries"):
        IA.Server(2048, 352, 32, max_queries=0)


def test_noalloc_scratch_sizes():
    # host-only sizing of inspire_respond_device_noalloc's per-GPU array: the
    # 1M-key and 32M-key shapes and the 4M-key ML-KEM-512 database (1024
    # outputs) fit with 2 hash queries; 100 hash queries over 64 outputs do not
    fits = [(2048, 352, 32, 2), (65536, 352, 32, 2), (2048, 1024, 32, 2), (2048, 8, 8, 1)]
    for rows, inst, d, nq in fits:
        need = IA.noalloc_scratch_bytes(rows, inst, d, nq)
        assert 0 < need <= IA.NOALLOC_SCRATCH_BYTES, (rows, inst, nq, need)
    assert IA.noalloc_scratch_bytes(2048, 64, 32, 100, max_queries=100) > IA.NOALLOC_SCRATCH_BYTES
    # per call: the packing and GEMV outputs dominate (about 64 KiB per output at nq 2)
    assert IA.noalloc_scratch_bytes(2048, 352, 32, 2) > 352 * 64 * 1024
    assert IA.noalloc_scratch_bytes(2048, 352, 32, 0) == 0
