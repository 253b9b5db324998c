"""Replay of recordings of the reference's own random draws.

Where MATLAB exists, radiativetransfermontecarlo_amd/matlab/recorder/
rtmc_record.m runs the reference's functions with every draw caught and
writes MAT files (tests/matlab_replay.py describes the layout); committed
under tests/golden/matlab/, they are replayed here through librtmc with
MATLAB's own prepared tables and T - mean(P.t): bins and counts must equal
MATLAB's, the particle states agree to the bars of the function-level
parity tests (times 1e-15 relative, positions and directions 1e-9; MATLAB's
sin/cos/acos and the device's may differ in the last bit).  No MATLAB
recording exists in this repository yet: those cases are skipped.

Recordings that the oracle writes in the same layout from synthetic draws
exercise the reader and the replay in the meantime: read back, the oracle
reproduces them bit for bit, and the GPU kernels agree with them."""
import pathlib

import numpy as np
import pytest

from tests import matlab_replay as mr

GOLDEN = pathlib.Path(__file__).parent / "golden" / "matlab"
RECORDINGS = sorted(GOLDEN.glob("*.mat")) if GOLDEN.is_dir() else []


def _check_states(got, want, exact):
    assert len(got) == len(want)
    for k, (g, w) in enumerate(zip(got, want)):
        for f in ('p', 'alive'):
            np.testing.assert_array_equal(np.asarray(g[f], dtype=np.float64), w[f], err_msg=f"call {k} {f}")
        a = w['alive'].astype(bool)
        if exact:
            for f in ('x', 'dir', 'perp', 't'):
                np.testing.assert_array_equal(g[f], w[f], err_msg=f"call {k} {f}")
            continue
        np.testing.assert_allclose(g['t'], w['t'], rtol=1e-15, atol=0, err_msg=f"call {k} t")
        scale = max(1.0, float(np.abs(w['x']).max()))
        np.testing.assert_allclose(g['x'], w['x'], rtol=0, atol=1e-9 * scale, err_msg=f"call {k} x")
        np.testing.assert_allclose(g['dir'][a], w['dir'][a], rtol=0, atol=1e-9, err_msg=f"call {k} dir")
        np.testing.assert_allclose(g['perp'][a], w['perp'][a], rtol=0, atol=1e-9, err_msg=f"call {k} perp")


def _recorded(rec):
    return [rec['init_state']] + rec['states'], np.asarray(rec['E']).ravel(order='F').astype(np.uint32)


@pytest.mark.parametrize("name,N", [("readme3d", 1500), ("halfspace3d", 1200), ("events3d", 1000)])
def test_synthetic_recording_round_trip_on_the_oracle(tmp_path, name, N):
    """Written in the recorder's layout and read back, a recording replays on
    the oracle to the same states and histogram bit for bit."""
    rec = mr.synthetic_recording(name, N, seed=11)
    path = tmp_path / f"{name}.mat"
    mr.save_recording(rec, path)
    back = mr.load_recording(path)
    assert back['kind'] == rec['kind'] and back['N'] == N and len(back['draws']) == len(rec['draws'])
    states, E = mr.replay_oracle(back)
    want_states, want_E = _recorded(back)
    _check_states(states, want_states, exact=True)
    np.testing.assert_array_equal(E, want_E)
    assert int(want_E.sum()) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name,N", [("readme3d", 1500), ("halfspace3d", 1200), ("events3d", 1000)])
def test_synthetic_recording_replays_on_gpu(gpu, tmp_path, name, N):
    rec = mr.synthetic_recording(name, N, seed=12)
    path = tmp_path / f"{name}.mat"
    mr.save_recording(rec, path)
    back = mr.load_recording(path)
    states, E = mr.replay_gpu(back, gpu)
    want_states, want_E = _recorded(back)
    _check_states(states, want_states, exact=False)
    np.testing.assert_array_equal(E, want_E)


@pytest.mark.gpu
@pytest.mark.skipif(not RECORDINGS, reason="no MATLAB recording: run radiativetransfermontecarlo_amd/matlab/recorder/"
                                            "rtmc_record.m where MATLAB exists and commit its files under "
                                            "tests/golden/matlab/")
@pytest.mark.parametrize("path", RECORDINGS, ids=[p.stem for p in RECORDINGS])
def test_matlab_recording_replays_exactly(gpu, path):
    rec = mr.load_recording(path)
    states, E = mr.replay_gpu(rec, gpu)
    want_states, want_E = _recorded(rec)
    np.testing.assert_array_equal(E, want_E)
    _check_states(states, want_states, exact=False)


REFERENCE = pathlib.Path("/root/reference")
RECORDER = pathlib.Path(__file__).parents[1] / "radiativetransfermontecarlo_amd" / "matlab" / "recorder" / "rtmc_record.m"
_FILES = {'initializeParticle': 'initializeParticle.m', 'propagateParticleSmallDt': 'propagateParticleSmallDt.m',
          'propagateParticle': 'propagateParticle.m', 'MaterialClass': '@MaterialClass/MaterialClass.m'}


def _recorder_sites():
    """(file, text, occurrence, slot) of every row of rtmc_record.m's site table."""
    import re
    lit = r"'((?:[^']|'')*)'"
    row = re.compile(r"^\s*" + lit + r",\s*" + lit + r",\s*(\d+),\s*" + lit + r",\s*(\d+),")
    out = []
    for line in RECORDER.read_text().splitlines():
        m = row.match(line)
        if m:
            out.append((m.group(1), m.group(2).replace("''", "'"), int(m.group(3)), int(m.group(5))))
    return out


@pytest.mark.skipif(not REFERENCE.is_dir(), reason="the reference checkout is not present")
def test_recorder_sites_cover_every_draw_of_the_reference():
    """The recorder's site table names every rand/randn/randi/poissrnd call of
    initializeParticle.m, propagateParticleSmallDt.m, propagateParticle.m and
    MaterialClass.m, each text at as many code lines as its occurrences (so
    the recorder finds each site, and only it, when it locates them)."""
    import re
    sites = _recorder_sites()
    assert len(sites) == 40
    covered = set()
    for f, text, occ, slot in sites:
        lines = (REFERENCE / _FILES[f]).read_text().splitlines()
        hits = [i for i, l in enumerate(lines) if text in l.strip()]
        assert len(hits) >= occ, (f, text)
        assert all(not lines[i].strip().startswith('%') for i in hits), (f, text)
        covered.add((f, hits[occ - 1]))
        assert 0 <= slot <= 7
    call = re.compile(r"\b(rand|randn|randi|poissrnd)\(")
    for f, rel in _FILES.items():
        for i, l in enumerate((REFERENCE / rel).read_text().splitlines()):
            if call.search(l.split('%')[0]):
                assert (f, i) in covered, (rel, i + 1, l.strip())
