"""CPU hardening under AddressSanitizer and UndefinedBehaviorSanitizer (every report fatal).

- The `.clusters` reader (host code that parses untrusted bytes), compiled from its own source into
  tests/sanitize/clusters_driver.cpp: a mutation fuzz of ge_clusters_read from two valid containers
  (content hash repaired for most mutants, so the section checks are reached), every read scene
  walked end to end; and every host allocation of a read failed in turn, each of which must come
  back as GE_RT_ERROR_OUT_OF_MEMORY with nothing leaked.
- The oracle built with both sanitizers (`make -C oracle asan`) under the CPU suite, run in a child
  Python with libasan preloaded (the brute-force walks, minutes under the sanitizers, left out).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from game_engine_amd import synth
from oracle import binding as ob
from tests import clusters_writer as cw

HERE = os.path.dirname(os.path.abspath(__file__))
SANITIZE = os.path.join(HERE, "sanitize")
LIBASAN = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()

pytestmark = pytest.mark.skipif(not os.path.isabs(LIBASAN), reason="no AddressSanitizer runtime for g++ here")
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0:exitcode=99",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-C", SANITIZE], check=True)
    return os.path.join(SANITIZE, "_build", "clusters_driver")


@pytest.fixture(scope="module")
def seeds(tmp_path_factory):
    # A small mesh (300 triangles in 3 clusters), so that a mutant costs little.
    rng = np.random.default_rng(4)
    base = rng.uniform(-1, 1, (300, 1, 3))
    s = synth.Synth.from_triangles(base + rng.normal(0, 0.1, (300, 3, 3)), cluster_size=128)
    m = np.zeros(2, dtype=cw.MATERIAL_DTYPE)
    m["base_color"] = [0.5, 0.5, 0.5, 0.75]
    m["base_color_image"], m["normal_image"] = 1, -1
    m[0]["alpha"] = 1 | (32768 << 16)
    cm = np.arange(s.desc().cluster_count) % 2
    d = tmp_path_factory.mktemp("clusters")
    paths = []
    for name, kw in (("quantized", dict(quantized=True, materials=m, cluster_material=cm)),
                     ("two_levels", dict(levels=2, quantized=False))):
        p = d / f"{name}.clusters"
        p.write_bytes(bytes(cw.from_synth(s, **kw)))
        paths.append(str(p))
    return paths


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env={**os.environ, **SAN_ENV, **(env or {})})
    print(r.stdout[-2000:], r.stderr[-4000:], flush=True)
    return r


def test_the_driver_is_instrumented(driver):
    r = _run([driver, "selfcheck"])
    assert r.returncode == 99 and "heap-buffer-overflow" in r.stderr


def test_clusters_reader_fuzz(driver, seeds):
    for i, seed in enumerate(seeds):
        r = _run([driver, "fuzz", seed, "20000", str(11 + i)])
        assert r.returncode == 0 and "fuzz: 20000 mutants" in r.stdout


def test_clusters_reader_allocation_failures(driver, seeds):
    for seed in seeds:
        r = _run([driver, "oom", seed])
        assert r.returncode == 0 and "each GE_RT_ERROR_OUT_OF_MEMORY; then read" in r.stdout


def test_cpu_suite_with_the_sanitized_oracle():
    subprocess.run(["make", "-s", "-C", os.path.dirname(ob.__file__), "asan"], check=True)
    lib = os.path.join(os.path.dirname(ob.__file__), "_build", "asan", "libge_oracle.so")
    r = _run([sys.executable, "-m", "pytest", "tests", "-m", "not gpu", "-q", "-x", "-p", "no:cacheprovider",
              "--ignore", os.path.join("tests", os.path.basename(__file__)),
              "-k", "not brute_force and not wide_walk_is_the_binary_walk"],
             env={"LD_PRELOAD": LIBASAN, "ASAN_OPTIONS": "detect_leaks=0:exitcode=99", "GE_ORACLE_LIB": lib}, timeout=900)
    assert r.returncode == 0 and " passed" in r.stdout
