"""CPU: the row-quad helpers of dec_device.h (a lane's four consecutive rows
loaded as whole dwords and decoded per row) built for the host with
AddressSanitizer and UBSan and checked against the per-row decoder
load_compact, for every row width 1..32 bytes and every partial last quad."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ultraprecise_amd", "csrc")


def test_quad_helpers_match_per_row_decoder(tmp_path):
    exe = str(tmp_path / "quad_host")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I" + CSRC,
                           os.path.join(HERE, "quad_host", "quad_host.cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "all checks passed" in r.stdout
