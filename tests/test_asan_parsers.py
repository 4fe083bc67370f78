"""AddressSanitizer run of libpirgpu's host-side parsers (CPU only).

The two parser sources, loader.cpp and proto_loader.cpp, are compiled for
the host alone (hipcc --cuda-host-only, AddressSanitizer on the host code)
and linked by g++ with the harness tests/asan/parser_fuzz.cpp ahead of the
built libpirgpu.so (the executable's definitions of the parsers' entry
points take precedence over the library's); the harness feeds truncated and mutated
inputs to pirg_server_load_from_buffer (a PirServer::SaveToFile image) and
pirg_server_create_from_proto (a serialized PirPreprocessedData). Without a
GPU an input that parses stops at the device check, so only the parsers
run: every input must be refused with an error code or reach that check,
and no read may leave an input (AddressSanitizer ends the run otherwise).
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import inspire_amd as ia
import oracle as O
import pir_proto
from helpers import PIR_SERVER_TEST_TYPES, pir_server_test_params, to_gpu_params
from inspire_amd import build as B

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIP_LIB = "/opt/rocm/lib"


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, " ".join(cmd) + "\n" + proc.stdout + proc.stderr
    return proc


def _build(tmp):
    inc = os.path.join(ROOT, "include")
    hip_flags = ["-x", "hip", "--cuda-host-only", f"--offload-arch={B.ARCH}", "-O1", "-g",
                 "-std=c++17", "-fPIC", "-ffp-contract=off", "-Xarch_host",
                 "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer", f"-I{inc}"]
    jobs, objs = [], []
    for src in ("loader.cpp", "proto_loader.cpp"):
        obj = os.path.join(tmp, src.replace(".", "_") + ".o")
        jobs.append([B.hipcc(), *hip_flags, "-c", os.path.join(B.CSRC, src), "-o", obj])
        objs.append(obj)
    obj = os.path.join(tmp, "parser_fuzz.o")
    jobs.append(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-fsanitize=address",
                 "-fno-omit-frame-pointer", f"-I{inc}", "-c",
                 os.path.join(HERE, "asan", "parser_fuzz.cpp"), "-o", obj])
    objs.append(obj)
    with ThreadPoolExecutor(max_workers=3) as ex:
        list(ex.map(_run, jobs))
    exe = os.path.join(tmp, "parser_fuzz")
    libdir = os.path.dirname(B.LIB_PATH)
    _run(["g++", "-fsanitize=address", *objs, B.LIB_PATH, f"-L{HIP_LIB}", "-lamdhip64",
          "-lpthread", f"-Wl,-rpath,{libdir}:{HIP_LIB}", "-o", exe])
    return exe


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(
    os.path.join(HIP_LIB, "libamdhip64.so")) or not os.path.exists(B.LIB_PATH),
    reason="needs g++, the HIP runtime and the built library")
def test_parsers_survive_truncated_and_mutated_inputs(tmp_path):
    if os.path.exists("/dev/kfd"):
        pytest.skip("with a GPU present, parsed inputs would go on to the device")
    # q = 2^62, two interpolation slices (so the second-dimension records
    # are present) and two shards, on a small DB: a 190 KB image, 570 KB proto.
    lq, q1, q2, _, mb = PIR_SERVER_TEST_TYPES["Uint64Mod62"]
    P = pir_server_test_params(lq, q1, q2, 2, mb, ell=2, n=256)
    s = O.sizes(P)
    db = O.fill_db_u16(5, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=3)
    (tmp_path / "params.bin").write_bytes(bytes(to_gpu_params(P).c()))
    (tmp_path / "image.bin").write_bytes(O.save_server(P, mat))
    (tmp_path / "proto.bin").write_bytes(pir_proto.serialize_data(P, mat))
    exe = _build(str(tmp_path))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    proc = subprocess.run([exe, str(tmp_path / "params.bin"), str(tmp_path / "image.bin"),
                           str(tmp_path / "proto.bin"), "7", "3000"],
                          capture_output=True, text=True, env=env, timeout=600)
    out = proc.stdout + proc.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert proc.returncode == 0, out[-4000:]
    for name in ("load_from_buffer", "create_from_proto"):
        line = re.search(rf"^{name}: (\d+) calls(.*) created=0$", proc.stdout, re.M)
        assert line, proc.stdout
        codes = dict((int(c), int(n)) for c, n in re.findall(r"rc(\d+)=(\d+)", line.group(2)))
        # The intact input (and mutations of bytes no check reads, such as DB
        # entries) parse and stop at the device check; the rest are refused.
        assert codes.get(ia.UNAVAILABLE, 0) >= 1, line.group(0)
        assert sum(n for c, n in codes.items() if c != ia.UNAVAILABLE) >= 1000, line.group(0)
        assert 0 not in codes
