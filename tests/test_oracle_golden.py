"""The oracle against its own committed history (tests/golden/config1_frames.json, written by
tests/golden/make_config1_frames.py): the small plain-DDGI workload's 10 frames must reproduce
the stored hit-record and state digests exactly and the stored atlas values to 1e-12."""
import json
import os

import numpy as np

from tests.golden.make_config1_frames import OUT, digest, run_oracle


def test_config1_frames_reproduce_the_fixture():
    want = json.load(open(OUT))
    frames = run_oracle()
    assert len(frames) == len(want["frames"])
    ip, dp = np.array(want["irradiance_positions"]), np.array(want["depth_positions"])
    for f, ((hits, states, irr, dep), w) in enumerate(zip(frames, want["frames"])):
        assert digest(hits) == w["hits_sha256"], f"frame {f}: hit records changed"
        assert digest(states) == w["states_sha256"], f"frame {f}: states changed"
        if "irradiance_samples" in w:
            assert np.allclose(irr.reshape(-1)[ip], w["irradiance_samples"], rtol=1e-12, atol=0)
            assert np.allclose(dep.reshape(-1)[dp], w["depth_samples"], rtol=1e-12, atol=0)
            assert np.isclose(irr.sum(), w["irradiance_sum"], rtol=1e-12)
            assert np.isclose(dep.sum(), w["depth_sum"], rtol=1e-12)
