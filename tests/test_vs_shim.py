"""The VapourSynth shim (vapoursynth-feel_amd/csrc/vsfeel_plugin.cpp) built on
the CPU against a hand-written declaration of the API4 types it touches
(tests/vs_shim/VapourSynth4.h) and driven by a fake core
(tests/vs_shim/fake_core.cpp):

* registration: the reference's identifier, namespace, function name and
  byte-exact argument string (src/vsfeel.cpp:350-364, src/bm3d.cpp:1558-1574);
* create errors, with the real library: mapSetError with "BM3D: ..." and the
  clip and ref nodes freed (src/bm3d.cpp:1043-1060);
* with a stand-in for the library (tests/vs_shim/stub_abi.cpp): the
  arguments reaching vsfeel_bm3d_create, the frame window requested, the
  planes and strides handed to vsfeel_bm3d_get_frame, chroma and properties
  copied for YUV, a failed request reported through setFilterError with its
  frame freed, and free releasing the handle and both nodes.
"""

from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "vs_shim")
PLUGIN = os.path.join(ROOT, "vapoursynth-feel_amd", "csrc", "vsfeel_plugin.cpp")
LIBDIR = os.path.join(ROOT, "vapoursynth-feel_amd", "vsfeel")

REF_ARGS = ("clip:vnode;ref:vnode:opt;sigma:float[]:opt;block_step:int[]:opt;bm_range:int[]:opt;"
            "radius:int:opt;ps_num:int[]:opt;ps_range:int[]:opt;num_streams:int:opt;extractor_exp:int:opt;")


def _compile(out, *extra):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + SHIM, "-I" + os.path.join(ROOT, "include"),
           PLUGIN, os.path.join(SHIM, "fake_core.cpp"), "-o", str(out), *extra]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return str(out)


@pytest.fixture(scope="module")
def real_core(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libvsfeel_bm3d.so")):
        pytest.skip("libvsfeel_bm3d.so is not built")
    return _compile(tmp_path_factory.mktemp("shim") / "real", "-L" + LIBDIR, "-lvsfeel_bm3d",
                    "-Wl,-rpath," + LIBDIR)


@pytest.fixture(scope="module")
def stub_core(tmp_path_factory):
    return _compile(tmp_path_factory.mktemp("shim") / "stub", "-DVSFEEL_STUB", os.path.join(SHIM, "stub_abi.cpp"))


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(": ")
        out[k] = v
    return out


def test_registration_matches_the_reference(stub_core):
    o = _run(stub_core, "register")
    assert o["identifier"] == "com.thefeeltrain.vsfeel"
    assert o["namespace"] == "vsfeel"
    assert o["function"] == "BM3Dv2"
    assert o["args"] == REF_ARGS
    assert o["return"] == "clip:vnode;"


@pytest.mark.parametrize("case,msg,ref_frees", [
    ("bad_sigma", 'BM3D: "sigma" must be non-negative', "-1"),
    ("bad_ref", 'BM3D: "ref" must be of the same dimensions as "clip"', "1"),
    ("bad_radius_ref", 'BM3D: "radius" must be in range [0, 4]', "1"),
])
def test_create_errors_free_the_nodes(real_core, case, msg, ref_frees):
    o = _run(real_core, "create", case)
    assert o["error"] == msg
    assert o["filter"] == ""
    assert o["clip_frees"] == "1" and o["ref_frees"] == ref_frees


def test_valid_arguments_with_the_real_library(real_core):
    # without a GPU creation fails at the device stage, through the same path
    o = _run(real_core, "create", "valid_ref")
    if o["error"]:
        assert o["error"].startswith("BM3D: ") and o["filter"] == ""
        assert o["clip_frees"] == "1" and o["ref_frees"] == "1"
    else:
        assert o["filter"] == "BM3D"


@pytest.mark.parametrize("case", ["gray", "yuv"])
def test_frame_requests_through_the_shim(stub_core, case):
    o = _run(stub_core, "frames", case)
    assert o["error"] == "" and o["filter"] == "BM3D"
    assert o["clip_frees"] == "0" and o["ref_frees"] == "0"
    assert o["mode_parallel"] == "1" and o["num_deps"] == "2"
    assert o["deps_nodes"] == "1" and o["deps_strict_spatial"] == "1"
    # the map's arrays reach the C ABI unchanged
    assert o["stub_sigma"] == "3 2 1"
    assert o["stub_ints"] == "4 16,9 1 2 7 3 2"
    cf = "3" if case == "yuv" else "1"
    assert o["stub_format"] == f"{cf} 1 32 {'1' if case == 'yuv' else '0'} 40 32 6 1"
    assert o["stub_modes"] == "1 0 0"  # fused multiply-add mode, device 0, plain FP32 form
    # frame 1 at radius 1: frames 0-3 of clip and ref (src/bm3d.cpp:704-712)
    assert o["initial_returns_null"] == "1"
    assert o["requested"].split() == ["0c", "0r", "1c", "1r", "2c", "2r", "3c", "3r"]
    assert o["frame_error"] == "" and o["stub_n"] == "1"
    assert o["planes_ok"] == "1"
    assert o["dst_luma_ok"] == "1" and o["dst_padding_untouched"] == "1"
    assert o["chroma_copied"] == "1" and o["props_ok"] == "1"
    assert o["live_frames_after_request"] == "0"
    assert o["stub_frees"] == "1" and o["live_handles"] == "0"
    assert o["clip_frees_after_free"] == "1" and o["ref_frees_after_free"] == "1"


def test_failed_request_sets_the_filter_error(stub_core):
    o = _run(stub_core, "frames", "fail")
    assert o["frame_error"] == "BM3D: stub launch failure"
    assert o.get("dst") == "null"
    assert o["live_frames_after_request"] == "0"  # the destination frame was freed
    assert o["stub_frees"] == "1" and o["clip_frees_after_free"] == "1"
