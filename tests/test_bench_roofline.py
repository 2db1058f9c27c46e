"""The bench's dominant-kernel roofline (host logic only, no GPU): the two
sweeps of one kernel function are taken together, so the pick does not flip
when one sweep alone is about as long per step as another kernel's two
launches."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def times(per_launch_us, launches):
    """{label: (total ms, launches)} as the library reports them."""
    return {k: (us * launches[k] * 1e-3, launches[k]) for k, us in per_launch_us.items()}


# per-launch HIP-event averages of a river run (20 steps): one sweep alone
# (25.65 us) is within 4 % of the state copy's two launches (2 x 12.3 us)
US = {"boundary": 6.09, "adi_spec_uz": 25.65, "adi_spec_vz": 25.29, "uvz_segments": 12.3,
      "smooth_u_explicit_v": 12.15, "smooth_v_explicit_u": 13.17}
N = {"boundary": 40, "adi_spec_uz": 20, "adi_spec_vz": 20, "uvz_segments": 40,
     "smooth_u_explicit_v": 20, "smooth_v_explicit_u": 20}


def test_sweeps_counted_together(tmp_path):
    tf = tmp_path / "traffic.json"
    tf.write_text(json.dumps({"adi_spec_uz": 12599157, "adi_spec_vz": 11487132, "uvz_segments": 22006867}))
    r = bench.roofline(times(US, N), 85556.3, str(tf))
    assert r["kernel"] == "adi_spec"
    assert r["labels"] == ["adi_spec_uz", "adi_spec_vz"]
    assert r["launches"] == 40
    assert r["avg_launch_us"] == pytest.approx((25.65 + 25.29) / 2)
    assert r["bytes_per_wet_cell"] == 88
    # achieved = 88 B x wet cells / average launch; frac against 8 TB/s
    assert r["achieved"] == pytest.approx(88 * 85556.3 / (r["avg_launch_us"] * 1e-6) / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["traffic"] == round((12599157 + 11487132) / 2)


def test_pick_does_not_flip_with_one_slow_state_copy(tmp_path):
    us = dict(US, uvz_segments=13.2)  # 26.4 us per step: above one sweep, below the pair
    r = bench.roofline(times(us, N), 85556.3, str(tmp_path / "none.json"))
    assert r["kernel"] == "adi_spec"
    assert r["traffic"] is None


def test_single_label_kernel_and_unknown_bytes(tmp_path):
    kt = times({"uvz_segments": 30.0, "adi_spec_uz": 20.0, "mystery": 5.0},
               {"uvz_segments": 40, "adi_spec_uz": 20, "mystery": 20})
    r = bench.roofline(kt, 1000.0, str(tmp_path / "none.json"))
    assert r["kernel"] == "uvz_segments" and r["labels"] == ["uvz_segments"]
    assert r["bytes_per_wet_cell"] == 84
    kt = times({"mystery": 50.0, "adi_spec_uz": 20.0}, {"mystery": 20, "adi_spec_uz": 20})
    r = bench.roofline(kt, 1000.0, str(tmp_path / "none.json"))
    assert r["kernel"] == "mystery" and r["achieved"] is None and r["frac"] is None
