"""The gather pins catch mistakes: each mutation below changes one token of the
oracle's per-pixel gather (oracle/gi_oracle.cpp gatherPixel, the restatement
of vpl_pbr_gi.fs:81-207), the mutated oracle is built into a temporary
directory with the oracle's own compiler flags, and at least one of the checks
of test_oracle_gather_pins.py must fail on it.  CPU only; the mistakes are
planted in copies of the oracle, never in the product."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle as orc
import test_oracle_gather_pins as pins

SRC = os.path.join(os.path.dirname(orc.__file__), "gi_oracle.cpp")
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-shared"]

# (name, text in gatherPixel, replacement); each text occurs exactly once in the file
MUTATIONS = [
    ("distance ratio 2d/R -> d/R", "clampv((R(2) * distance) / lightRadius", "clampv((R(1) * distance) / lightRadius"),
    ("occlusion floor mix ends swapped", "mixv(minGiOcc, R(1), distanceRatio)", "mixv(R(1), minGiOcc, distanceRatio)"),
    ("reweighting taken at shadow >= 0", "if (shadowFactor > R(0)) reweightingFactor",
     "if (shadowFactor >= R(0)) reweightingFactor"),
    ("reweighting minus the ratio", "* minGiOcc + distAttenuation;", "* minGiOcc - distAttenuation;"),
    ("SSAO floor 0.35 -> 0.25", "R(0.35), R(1));", "R(0.25), R(1));"),
    ("shadow cutoff 500 -> 5000", "distToCamera < R(500)", "distToCamera < R(5000)"),
    ("shadow min bias 0.1 -> 0.01", "R(0.1))\n                             : R(0);", "R(0.01))\n                             : R(0);"),
    ("sample count mix(2, 10) ends swapped", "mixv(R(2), R(10), roughnessWeight)", "mixv(R(10), R(2), roughnessWeight)"),
    ("camera distance scale 1000 -> 100", "distToCamera / R(1000)", "distToCamera / R(100)"),
    ("camera ratio min -> max", "gmin(R(1) - distToCamera / R(1000), R(1))", "gmax(R(1) - distToCamera / R(1000), R(1))"),
    ("at least one sample -> zero", "std::max(1, (int)(R(samplesMax) * distRatioToCamera))",
     "std::max(0, (int)(R(samplesMax) * distRatioToCamera))"),
    ("roughness floor 0.5 -> 0.05", "roughness = gmax(R(0.5), roughness);", "roughness = gmax(R(0.05), roughness);"),
    ("slot int((N-1) rand) -> int(N rand)", "(int)(R(maxRandomIndex) * R(rnd))", "(int)(R(numVisible) * R(rnd))"),
    ("rejection limit 5 -> 4", "evalVpl(lightIndex, resamples < 5)", "evalVpl(lightIndex, resamples < 4)"),
    ("seed step 10000 -> 1000 (reference estimator)",
     "seedZ += 10000.0f;\n            const int lightIndex = (int)(R(maxRandomIndex)",
     "seedZ += 1000.0f;\n            const int lightIndex = (int)(R(maxRandomIndex)"),
    ("tile list read from slot 0", "evalVpl(tileList[1 + k], true)", "evalVpl(tileList[k], true)"),
    ("side test sign flipped", "if (sideCheck < R(0) || distance > lightRadius)",
     "if (sideCheck > R(0) || distance > lightRadius)"),
    ("side or radius -> side and radius", "if (sideCheck < R(0) || distance > lightRadius)",
     "if (sideCheck < R(0) && distance > lightRadius)"),
    ("light times (1 + shadow)", "vplColor + (R(1) - shadowFactor) * tmpColor", "vplColor + (R(1) + shadowFactor) * tmpColor"),
    ("weight floor 1 -> 0", "validSamples = gmax(validSamples, R(1));", "validSamples = gmax(validSamples, R(0));"),
    ("albedo floor 0.01 -> 0.001", "gmax(baseColor.x, R(0.01))", "gmax(baseColor.x, R(0.001))"),
    ("HDR bound 65504 -> 65505", "vmin(vplColor, R(65504))", "vmin(vplColor, R(65505))"),
    ("AO replaced by the ambient channel", "ambientOcclusion * calculateVirtualPointLighting2<R>",
     "ambient * calculateVirtualPointLighting2<R>"),
    ("specular / diffuse positions swapped", "specularLightPosition, lightPosition, lightColor",
     "lightPosition, specularLightPosition, lightColor"),
    ("roughness / metallic swapped", "distToCamera, lightRadius, roughness, metallic,",
     "distToCamera, lightRadius, metallic, roughness,"),
    ("fixed-count slot int(N rand) -> int((N-1) rand)", "(int)((float)numVisible * rnd)",
     "(int)((float)(numVisible - 1) * rnd)"),
]


def _build(src_text: str, name: str, outdir: str) -> str:
    d = os.path.join(outdir, name)
    os.makedirs(d)
    cpp = os.path.join(d, "gi_oracle.cpp")
    with open(cpp, "w") as f:
        f.write(src_text)
    so = os.path.join(d, "liboracle.so")
    # "../include/sgi.h" resolves through -I against the oracle's own directory
    subprocess.run(["g++", *CXXFLAGS, "-I", os.path.dirname(SRC), "-o", so, cpp], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    return so


@pytest.fixture(scope="module")
def mutated_builds():
    with open(SRC) as f:
        text = f.read()
    lo, hi = text.index("void gatherPixel("), text.index("// ---- direct light")
    jobs = []
    for i, (name, old, new) in enumerate(MUTATIONS):
        assert text.count(old) == 1, f"mutation text not unique: {name}"
        at = text.index(old)
        assert lo <= at < hi, f"mutation outside gatherPixel: {name}"
        jobs.append((f"m{i:02d}", text.replace(old, new)))
    tmp = tempfile.mkdtemp(prefix="oracle_mut_")
    try:
        with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
            libs = list(ex.map(lambda j: _build(j[1], j[0], tmp), jobs))
        yield libs
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_unmutated_copy_passes_every_check(tmp_path):
    """The harness itself: a copy built the same way passes every pin."""
    with open(SRC) as f:
        so = _build(f.read(), "copy", str(tmp_path))
    L = orc.load(so)
    for check in pins.CHECKS:
        check(L)


@pytest.mark.parametrize("idx", range(len(MUTATIONS)), ids=[m[0] for m in MUTATIONS])
def test_mutation_is_caught(mutated_builds, idx):
    L = orc.load(mutated_builds[idx])
    failed = []
    for check in pins.CHECKS:
        try:
            check(L)
        except AssertionError:
            failed.append(check.__name__)
    assert failed, f"no gather pin catches: {MUTATIONS[idx][0]}"
