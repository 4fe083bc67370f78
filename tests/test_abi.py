"""The C-ABI boundary without a GPU: both libraries load, every function the
headers declare is exported, and PirServer::Create's validation
(crypto/pir/server/pir_server.cc:62-95, pir_params.cc:35-62) returns the
reference's status codes and messages before any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from tests.helpers import PIR_SERVER_TEST_TYPES, pir_server_test_params, to_gpu_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    text = open(header).read()
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, text)))


def test_libpirgpu_exports_every_declared_symbol():
    names = _declared(os.path.join(ROOT, "include", "pirgpu.h"), "pirg_")
    assert len(names) >= 10
    lib = ia.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_probes_library_exports_its_header_and_the_product_does_not():
    # The calibration probes live in their own library, outside the answer
    # path's ABI.
    from inspire_amd import probes
    names = _declared(os.path.join(ROOT, "include", "pirgpu_probes.h"), "pirg_")
    assert names == ["pirg_copy_probe", "pirg_mfma_probe", "pirg_mfma_probe_random",
                     "pirg_read_probe"]
    plib = probes.lib()
    assert all(hasattr(plib, n) for n in names)
    product = set(_declared(os.path.join(ROOT, "include", "pirgpu.h"), "pirg_"))
    assert not product & set(names)


def test_liboracle_exports_every_declared_symbol():
    names = _declared(os.path.join(ROOT, "oracle", "pir_oracle.h"), "orc_")
    lib = O.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert "gfx950" in ia.version()


def _view_args(P, **over):
    s = O.sizes(P)
    ell, t, d, k, e = s.ell, s.t, s.d, s.k, s.e
    args = dict(
        db_matrices=[np.zeros((t * d, s.K // 4), np.uint64) for _ in range(ell)],
        db_cols=s.K,
        pack=[[np.zeros((d, k * d // 2), np.uint64) for _ in range(t)] for _ in range(ell)],
        a_tilde=[[np.zeros(d, np.uint64) for _ in range(t)] for _ in range(ell)],
        t_vec_h=[[np.zeros((k, d), np.uint64) for _ in range(t)] for _ in range(ell)],
        second_dim_a=np.zeros((2 * e, d), np.uint64),
    )
    args.update(over)
    return args


@pytest.mark.parametrize("case,msg", [
    ("num_db", "db_matrices size mismatch."),
    ("rows", "DB matrix must have td rows."),
    ("cols", "DB matrix must have at least N/t columns."),
    ("cols_differ", "All DB matrices must have the same number of columns."),
    ("outputs", "preprocessed_outputs size mismatch."),
    ("outputs_len", "Preprocessed outputs must be of length t."),
    ("sda", "second_dim_query_a must have length 2 * PolyEvalGadgetParams().num_digits."),
])
def test_create_validation_messages(case, msg):
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"], ell=2)
    s = O.sizes(P)
    a = _view_args(P)
    kw = {}
    if case == "num_db":
        a["db_matrices"] = a["db_matrices"][:1]
    elif case == "rows":
        kw["db_rows"] = [1, s.t * s.d]
    elif case == "cols":
        kw["db_cols_each"] = [1, 1]
    elif case == "cols_differ":
        kw["db_cols_each"] = [s.K, s.K + 4]
    elif case == "outputs":
        for key in ("pack", "a_tilde", "t_vec_h"):
            a[key] = a[key][:1]
    elif case == "outputs_len":
        for key in ("pack", "a_tilde", "t_vec_h"):
            a[key][0] = a[key][0][:-1]
    elif case == "sda":
        a["second_dim_a"] = a["second_dim_a"][:-1]
    if case == "cols":
        a["db_cols"] = 1
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create(to_gpu_params(P), a["db_matrices"], a["db_cols"], a["pack"],
                            a["a_tilde"], a["t_vec_h"], a["second_dim_a"], **kw)
    assert ei.value.code == ia.INVALID_ARGUMENT
    assert ei.value.message == msg


@pytest.mark.parametrize("over,msg", [
    (dict(num_entries=0), "num_entries must be strictly positive."),
    (dict(interp_degree=3), "interpolation_degree must be a power of two."),
    (dict(num_entries=1000), "interpolation_degree (32) must divide num_entries (1000)."),
    (dict(interp_degree=128),
     "interpolation_degree (128) must be less than or equal to 2 * ring degree (64)."),
])
def test_create_params_validation(over, msg):
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"])
    gp = to_gpu_params(P)
    gp = ia.PirParams(**{**{f: getattr(gp, f) for f in ia.PirParams.__dataclass_fields__},
                         **over})
    a = _view_args(P)
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create(gp, a["db_matrices"], a["db_cols"], a["pack"], a["a_tilde"],
                            a["t_vec_h"], a["second_dim_a"])
    assert ei.value.code == ia.INVALID_ARGUMENT
    assert ei.value.message == msg


def test_create_without_device_fails_loudly():
    """A valid create either succeeds (GPU present) or reports UNAVAILABLE;
    there is no CPU fallback."""
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62Interp1"])
    a = _view_args(P)
    try:
        srv = ia.PirServer.create(to_gpu_params(P), a["db_matrices"], a["db_cols"], a["pack"],
                                  a["a_tilde"], a["t_vec_h"], a["second_dim_a"])
        srv.close()
    except ia.PirError as e:
        assert e.code == ia.UNAVAILABLE


@pytest.mark.parametrize("n", [4, 64, 4096, 8192])
def test_fft_twiddles_match_oracle(n):
    """The FFT kernels' twiddle table is bit-identical to the oracle's FftPlan
    (crypto/fft.h:71-80) and within one ulp of cos/sin."""
    import math
    dp = ctypes.POINTER(ctypes.c_double)
    a, b, c, d = (np.zeros(n) for _ in range(4))
    assert ia.lib().pirg_fft_twiddles(n, a.ctypes.data_as(dp), b.ctypes.data_as(dp)) == 0
    O.lib().orc_fft_twiddles(n, c.ctypes.data_as(dp), d.ctypes.data_as(dp))
    np.testing.assert_array_equal(a, c)
    np.testing.assert_array_equal(b, d)
    for i in range(0, n // 2, max(1, n // 64)):
        ang = -2.0 * math.pi * i / n
        for got, want in [(a[2 * i], math.cos(ang)), (a[2 * i + 1], math.sin(ang)),
                          (b[2 * i], math.cos(-ang)), (b[2 * i + 1], math.sin(-ang))]:
            assert abs(got - want) <= 2 * math.ulp(max(abs(want), 1e-300))


@pytest.mark.parametrize("length,logm,nbytes,msg", [
    (3, 8, 3, "Invalid length."),
    (4, 0, 4, "Invalid log_modulus."),
    (4, 8, 0, "Empty serialized polynomial."),
    (4, 8, 5, "Unexpected serialized size."),
])
def test_poly_from_proto_checks_before_the_device(length, logm, nbytes, msg):
    # CreateFromProto's messages (crypto/polynomial.cc:211-226), no GPU needed.
    with pytest.raises(ia.PirError) as ei:
        ia.poly_from_proto([bytes(nbytes)], length, logm)
    assert ei.value.code == ia.INVALID_ARGUMENT and ei.value.message == msg


def test_poly_to_proto_rejects_bad_log_modulus():
    with pytest.raises(ia.PirError) as ei:
        ia.poly_to_proto(np.zeros((1, 4), np.uint64), 0)
    assert ei.value.message == "Invalid log_modulus."


def test_load_from_buffer_reads_and_checks_before_the_device():
    P = O.metrics_tool_params(n=2048, log_d=8, t=2)
    img = O.save_server(P, O.synthetic_material(P, seed=1))
    gp = to_gpu_params(P)
    # Cut inside the pack record: "Error reading from buffer" (INTERNAL), as
    # LoadFromBuffer returns when it stops short of the end.
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.load_from_buffer(gp, img[:len(img) // 2])
    assert ei.value.code == 13 and ei.value.message == "Error reading from buffer"
    # An image of other parameters.
    P2 = O.metrics_tool_params(n=2048, log_d=8, t=4)
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.load_from_buffer(to_gpu_params(P2), img)
    assert ei.value.code == ia.INVALID_ARGUMENT


def test_proto_loader_checks_before_the_device(tmp_path):
    import pir_proto
    P = O.metrics_tool_params(n=2048, log_d=8, t=2)
    gp = to_gpu_params(P)
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create_from_proto(gp, b"\xff\xff\xff")
    assert ei.value.message == "Failed to parse PirPreprocessedData from file."
    mat = O.synthetic_material(P, seed=1)
    msg = pir_proto.M["PirPreprocessedData"]()
    msg.ParseFromString(pir_proto.serialize_data(P, mat))
    # CreateFromProto's size check on a_tilde_agg (polynomial.cc:221-226).
    msg.preprocessed_outputs[0].outputs[0].a_tilde_agg.encoded_coeffs = b"\x01\x02"
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create_from_proto(gp, msg.SerializeToString())
    assert ei.value.message == "Unexpected serialized size."
    # Matrix::FromProto (crypto/matrix.cc:504-507).
    msg.ParseFromString(pir_proto.serialize_data(P, mat))
    m = msg.preprocessed_outputs[0].outputs[1].matrix
    m.data = m.condensed_data
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create_from_proto(gp, msg.SerializeToString())
    assert ei.value.message == "Condensed matrix cannot have normal data set."
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create_from_directory(gp, str(tmp_path / "missing"))
    assert ei.value.code == 5


# ---- Offline preprocessing: argument checks run before any device call ----

def test_preprocessing_stage_checks_before_the_device():
    rows = np.zeros(64, dtype=np.uint16)
    for args, msg in [((rows, 12, 2, 257), "ring_degree must be a power of two."),
                      ((rows, 16, 3, 257), "interpolation_degree must be a power of two."),
                      ((rows, 2, 8, 257), "less than or equal to 2 * ring_degree"),
                      ((rows, 16, 4, 256), "plaintext_modulus must be odd.")]:
        with pytest.raises(ia.PirError) as ei:
            ia.interpolate(*args)
        assert ei.value.code == ia.INVALID_ARGUMENT and msg in ei.value.message
    # t == 1 is the identity (db_row_interpolation.cc:141-143) and needs no device.
    np.testing.assert_array_equal(ia.interpolate(np.arange(8, dtype=np.uint16), 8, 1, 257),
                                  np.arange(8))
    a = np.zeros((1, 4), dtype=np.uint64)
    for db, comps, msg in [(np.zeros((0, 4), np.uint16), a, "Database matrix cannot be empty."),
                           (np.zeros((2, 4), np.uint16), np.zeros((0, 4), np.uint64),
                            "a_components cannot be empty."),
                           (np.zeros((2, 3), np.uint16), a, "Database columns must match d")]:
        with pytest.raises(ia.PirError) as ei:
            ia.compute_hint_matrix(db, comps)
        assert ei.value.code == ia.INVALID_ARGUMENT and msg in ei.value.message


def test_preprocess_checks_before_the_device():
    P = O.metrics_tool_params(n=512, log_d=8, t=2)
    s = O.sizes(P)
    db = np.zeros(P.num_entries * s.d, dtype=np.uint16)
    first, second, pack_a = O.sample_a_components(P, 1)
    bad = ia.PirParams(**{**{f: getattr(P, f) for f in ia.PirParams.__dataclass_fields__},
                          "plaintext_modulus": 65536})
    with pytest.raises(ia.PirError) as ei:
        ia.preprocess(bad, db, first, pack_a)
    assert ei.value.message == "plaintext_modulus must be odd."
    bad = ia.PirParams(**{**{f: getattr(P, f) for f in ia.PirParams.__dataclass_fields__},
                          "num_entries": 513})
    with pytest.raises(ia.PirError) as ei:
        ia.preprocess(bad, np.zeros(513 * s.d, np.uint16), first, pack_a)
    assert "must divide num_entries" in ei.value.message
    # Null arguments through the C ABI.
    err = ctypes.create_string_buffer(256)
    cp = to_gpu_params(P).c()
    rc = ia.lib().pirg_preprocess(ctypes.byref(cp), None, None, None, None, None, None, None,
                                  None, None, err, ctypes.c_size_t(len(err)))
    assert rc == ia.INVALID_ARGUMENT and b"must not be null" in err.value


def test_preprocess_without_device_fails_loudly():
    """Valid arguments either preprocess (GPU present) or report UNAVAILABLE;
    there is no CPU fallback."""
    P = O.metrics_tool_params(n=512, log_d=8, t=2)
    s = O.sizes(P)
    db = np.zeros(P.num_entries * s.d, dtype=np.uint16)
    first, second, pack_a = O.sample_a_components(P, 1)
    try:
        ia.preprocess(to_gpu_params(P), db, first, pack_a)
    except ia.PirError as e:
        assert e.code == ia.UNAVAILABLE


@pytest.mark.parametrize("over,msg", [
    (dict(coeff_bits=16), "coeff_bits must be 32 or 64."),
    (dict(coeff_bits=32), "log_q exceeds the coefficient width (32 bits)."),
    (dict(coeff_bits=32, log_q=32, log_q1=31, log_q2=30, mat_bytes=8),
     "mat_bytes 8 (a uint64_t pack matrix) needs 64-bit coefficients."),
])
def test_coefficient_type_validation(over, msg):
    # CoeffType is uint32_t or uint64_t (pir_server.cc:806-818); a uint32_t
    # modulus has at most 32 bits (RlweParams::LogModulus, encryption.h:46-49).
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"])
    gp = to_gpu_params(P)
    gp = ia.PirParams(**{**{f: getattr(gp, f) for f in ia.PirParams.__dataclass_fields__},
                         **over})
    a = _view_args(P)
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create(gp, a["db_matrices"], a["db_cols"], a["pack"], a["a_tilde"],
                            a["t_vec_h"], a["second_dim_a"])
    assert ei.value.code == ia.INVALID_ARGUMENT
    assert ei.value.message == msg


def test_uint64_pack_matrix_outside_int32_is_rejected_before_the_device():
    # PirServer<Db, uint64_t, uint64_t>: entries are held as int32 on the GPU,
    # which BuildPreprocessMatrix's centered values always fit.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62Interp1"])
    gp = ia.PirParams(**{**{f: getattr(to_gpu_params(P), f)
                            for f in ia.PirParams.__dataclass_fields__}, "mat_bytes": 8})
    s = O.sizes(P)
    a = _view_args(P)
    pack = [[np.zeros((s.d, s.k * s.d), np.uint64) for _ in range(s.t)] for _ in range(s.ell)]
    pack[0][0][3, 5] = np.uint64((1 << 64) - (1 << 31))      # -2^31: fits
    pack[0][0][4, 1] = np.uint64((1 << 64) - (1 << 62) - 16)  # -16 mod 2^62: fits
    pack[0][0][7, 9] = np.uint64(1 << 31)                     # 2^31: does not
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.create(gp, a["db_matrices"], a["db_cols"], pack, a["a_tilde"],
                            a["t_vec_h"], a["second_dim_a"])
    assert ei.value.code == ia.INVALID_ARGUMENT
    assert ei.value.message.startswith(f"pack matrix entry {7 * s.k * s.d + 9} of block 0 lies "
                                       "outside [-2^31, 2^31) mod q")


def test_entry_points_reject_null_and_negative_arguments():
    # Checked before any allocation or device call, so no GPU is needed.
    lib = ia.lib()
    err = ctypes.create_string_buffer(256)
    n = ctypes.c_size_t(256)
    out = ctypes.c_void_p()
    assert lib.pirg_server_create(None, None, ctypes.byref(out), err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_server_create_flat(None, None, 0, None, None, None, None,
                                       ctypes.byref(out), err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_process_batch(None, 1, None, None, None, err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_batch_stage(None, -1, None, err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_batch_run(None, 1, err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_batch_fetch(None, 1, None, err, n) == ia.INVALID_ARGUMENT
    fused = ctypes.c_int()
    assert lib.pirg_server_scans_fused(None, 64, ctypes.byref(fused)) == ia.INVALID_ARGUMENT
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62Interp1"])
    gp = ia.PirParams(**{**{f: getattr(to_gpu_params(P), f)
                            for f in ia.PirParams.__dataclass_fields__}, "log_degree": 40})
    cp = gp.c()
    rc = lib.pirg_vpir_server_create(ctypes.byref(cp), 0, None, 0, 0, None, 0, None, None,
                                     ctypes.byref(out), err, n)
    assert rc == ia.INVALID_ARGUMENT
    assert b"log_degree" in err.value


def test_record_submission_entry_points_reject_bad_arguments():
    lib = ia.lib()
    err = ctypes.create_string_buffer(256)
    n = ctypes.c_size_t(256)
    p = ctypes.c_void_p()
    assert lib.pirg_host_alloc(0, ctypes.byref(p), err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_host_alloc(64, None, err, n) == ia.INVALID_ARGUMENT
    lib.pirg_host_free(None)
    # A pointer pirg_host_alloc did not return is ignored, not freed.
    junk = ctypes.create_string_buffer(64)
    lib.pirg_host_free(ctypes.cast(junk, ctypes.c_void_p))
    assert lib.pirg_request_record_words(None) == 0
    assert lib.pirg_request_record(None, None, None, err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_serve_submit_records(None, 1, None, None, err, n) == ia.INVALID_ARGUMENT
    assert lib.pirg_serve_submit_records(None, -1, None, None, err, n) == ia.INVALID_ARGUMENT


def test_load_from_file_and_save_errors(tmp_path):
    # LoadFromFile's read failure (pir_server.cc:799-801) and SaveToFile's
    # open failure (:437-440), before any device work.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"])
    with pytest.raises(ia.PirError) as ei:
        ia.PirServer.load_from_file(to_gpu_params(P), str(tmp_path / "missing.bin"))
    assert ei.value.code == ia.INTERNAL
    assert ei.value.message == "Failed to read file."
    lib = ia.lib()
    err = ctypes.create_string_buffer(256)
    assert lib.pirg_server_save_to_file(None, b"/tmp/x", err, 256) == ia.INVALID_ARGUMENT
    size = ctypes.c_size_t()
    assert lib.pirg_server_save_to_buffer(None, None, 0, ctypes.byref(size), err,
                                          256) == ia.INVALID_ARGUMENT
