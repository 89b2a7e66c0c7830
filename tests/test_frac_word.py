"""The FP32 storage's colour word (porescale_amd/csrc/lbm_frac.h), built for
the host with g++ and checked against what rounding to a 25-bit significand
(33 bits with the extension byte) fixes: the nearest representable fraction,
its error bounds, exact single-colour patterns and a lossless decode/encode
round trip."""
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "porescale_amd", "csrc")


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("frac") / "frac_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "cpp", "frac_check.cpp"), "-o", exe],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rep = {}
    for line in out.splitlines():
        rep.update(json.loads(line))
    return rep


def test_fraction_rounds_to_nearest_25_bit(report):
    assert report["not_nearest"] == 0
    assert report["max_rel"] <= 2.0 ** -25
    assert report["max_abs_tiny"] <= 2.0 ** -151


def test_minority_part_keeps_its_relative_precision(report):
    # minority parts from 1e-30 to 1/2 of the population
    assert report["max_min_rel"] <= 2.0 ** -25 * (1 + 2.0 ** -20)
    assert report["max_maj"] <= 2.0 ** -26


def test_words_round_trip(report):
    assert report["roundtrip_words"] > 2_000_000
    assert report["roundtrip_bad"] == 0


def test_escape_for_parts_of_opposite_sign(report):
    assert report["esc_words"] == 6
    assert report["max_esc"] <= 2.0 ** -16
    assert report["near1_escape"] == 1 and report["near1_value"] == 1.0


def test_single_colour_patterns_are_exact(report):
    assert report["all_red"] == 1 and report["all_blue"] == 0
    assert report["red_pure"] == 1 and report["blue_pure"] == 1
    assert report["zero_split"] == [0.0, 0.0]


def test_extended_word_rounds_to_33_bits(report):
    assert report["max_half_ulp_x"] <= 0.5 * (1 + 1e-12)
    assert report["max_rel_x"] <= 2.0 ** -33
    assert report["max_abs_tiny_x"] <= 2.0 ** -159
    assert report["max_min_rel_x"] <= 2.0 ** -33 * (1 + 2.0 ** -12)
    assert report["roundtrip_words_x"] > 1_000_000 and report["roundtrip_bad_x"] == 0
    assert report["pure_x_nonzero"] == 0

