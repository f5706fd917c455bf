"""Host test of the serving loop's admission arena and FIFO rule
(quicktex_amd/csrc/server_core.h), compiled with g++ and run on the CPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_server_core(tmp_path):
    exe = str(tmp_path / 'server_core_test')
    # host-only build, with the address and undefined-behaviour sanitizers
    subprocess.run(['g++', '-std=c++17', '-O1', '-g', '-Wall', '-fsanitize=address,undefined',
                    '-I' + os.path.join(ROOT, 'quicktex_amd', 'csrc'),
                    os.path.join(ROOT, 'tests', 'server_core_test.cpp'), '-o', exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and 'server_core ok' in r.stdout, r.stdout + r.stderr
