"""CPU builds of the oracle and of the host emulation library with
AddressSanitizer and UndefinedBehaviorSanitizer, each driven through every
entry point (oracle/sanitize_check.c, tests/cpp/hostemu_sanitize.cpp).  The
sanitizers abort on an out-of-bounds access, use after free, leak,
misaligned access, signed overflow or invalid shift; the drivers also fail
on an index the emulation reports out of range.  (Host code only: GPU
sanitizers are not used.)"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wayland_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang++"


def _run(exe):
    # verify_asan_link_order=0: the environment may preload other libraries
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)


def test_oracle_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    p = _run(os.path.join(ROOT, "oracle", "_san", "oracle_check"))
    assert p.returncode == 0, p.stdout + p.stderr
    assert "SANITIZE_CHECK_OK" in p.stdout


def test_hostemu_under_asan_ubsan():
    out_dir = os.path.join(ROOT, "wayland_amd", "_build")
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, "hostemu_sanitize")
    cmd = [CLANG, "-x", "c++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(CSRC, "hostemu.cpp"), os.path.join(CSRC, "bvh_builder.cpp"),
           os.path.join(ROOT, "tests", "cpp", "hostemu_sanitize.cpp"),
           "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    p = _run(exe)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "HOSTEMU_SANITIZE_OK" in p.stdout
