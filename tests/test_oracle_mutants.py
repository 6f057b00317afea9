"""The oracle's shading pins bite: each planted mistake fails a named pin.

oracle/ddgi_oracle.c carries switches for eleven plausible mistakes, compiled in
only with -DORA_MUTANT=<id> (the default build has none).  This test builds each
mutant into a temporary directory with the flags of oracle/Makefile, runs
tests/test_oracle_pins.py against it in a child process (DDGI_ORACLE_LIB points
the oracle wrapper at the mutant) and checks that the pins named below fail.
"""
import os
import subprocess
import sys
import xml.etree.ElementTree as ET
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ddgi_oracle.c")
PINS = os.path.join(ROOT, "tests", "test_oracle_pins.py")
CFLAGS = ["-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]

DIRECT = ["test_direct_light_symmetric_view_and_light", "test_direct_light_asymmetric_view_and_light"]
INDIRECT = ["test_gi_lookup_fresnel_roughness_off_normal", "test_probe_hit_indirect_off_normal"]
MIDPOINTS = ["test_midpoint_between_two_probes_in_x", "test_midpoint_between_layers_and_in_z",
             "test_cell_centre_two_weight_groups"]
RAMPS = ["test_tile_texel_addressing_with_linear_ramps"]

# id: (mistake, pins that must fail)
MUTANTS = {
    1: ("Schlick-GGX k = r^2 / 2", DIRECT),
    2: ("Fresnel exponent 4", DIRECT + INDIRECT),
    3: ("wrap weight without +0.2", MIDPOINTS),
    4: ("corners (1,1,0) and (1,0,0) swapped", ["test_midpoint_between_two_probes_in_x",
                                                "test_grid_edges_wrap_as_the_sampler_repeats"]),
    5: ("tile start off by one", RAMPS),
    6: ("trilinear weights flipped", ["test_on_probe_reads_its_own_tile",
                                      "test_grid_edges_wrap_as_the_sampler_repeats"] + RAMPS),
    7: ("indirect Fresnel with roughness, not roughness^2", INDIRECT),
    8: ("tile end off by one", RAMPS),
    9: ("GGX numerator without +0.001", DIRECT),
    10: ("probes blended in linear space", MIDPOINTS),
    11: ("octEncode without the y/z swap", RAMPS),
}


def _run(mid, tmp):
    so = os.path.join(tmp, f"libddgi_oracle_m{mid}.so")
    subprocess.run(["gcc"] + CFLAGS + [f"-DORA_MUTANT={mid}", "-o", so, SRC, "-lm"], check=True)
    xml = os.path.join(tmp, f"m{mid}.xml")
    env = dict(os.environ, DDGI_ORACLE_LIB=so, OMP_NUM_THREADS="1")
    subprocess.run([sys.executable, "-m", "pytest", PINS, "-q", "-p", "no:cacheprovider",
                    f"--junitxml={xml}"], cwd=ROOT, env=env, capture_output=True, text=True)
    failed, total = set(), 0
    for case in ET.parse(xml).getroot().iter("testcase"):
        total += 1
        if case.find("failure") is not None or case.find("error") is not None:
            failed.add(case.get("name").split("[")[0])
    return failed, total


@pytest.fixture(scope="module")
def outcomes(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("mutants"))
    with ThreadPoolExecutor(max_workers=4) as ex:
        futs = {m: ex.submit(_run, m, tmp) for m in MUTANTS}
        return {m: f.result() for m, f in futs.items()}


@pytest.mark.parametrize("mid", sorted(MUTANTS))
def test_planted_mistake_fails_its_pins(outcomes, mid):
    failed, total = outcomes[mid]
    name, must_fail = MUTANTS[mid]
    assert total >= 14, f"mutant {mid}: pins did not run"
    missing = [t for t in must_fail if t not in failed]
    assert not missing, f"mutant {mid} ({name}) passed pins {missing}"


def test_unmutated_build_passes_every_pin(tmp_path):
    failed, total = _run(0, str(tmp_path))
    assert total >= 14 and not failed, failed
