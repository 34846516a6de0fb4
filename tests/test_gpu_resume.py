"""Checkpoint / resume of the wz_match-compatible CLI against the reference's own runs.

Each case of tests/golden/resume.json is an interrupt/resume pair run by the reference:
WZ_FH_TEST_STOP_AFTER=k fakes a SIGTERM after exactly k completions (wz_match.cpp:2608-2612),
the arm writes arm_<shard>.ckpt (CFGSIG, resume_pi/batch/k and the cumulative counters,
:2464-2477), the checkpoint is optionally forged, and a second run resumes from it
(:2288-2317, the batch- and rank-level fast-forward at :2656-2680). The CLI must print the same
result lines and write byte-identical checkpoints in both runs. The cases follow the
reference's own gates: mid-drain interrupts with the stream reversed and buffered or not
(tools/test_firsthit_telemetry.py:129-145), forged buffer-boundary checkpoints (:146-171),
front-only checkpoints forged to batch == DRAIN_BATCHES, k = 0, which must complete nothing
more in that cell (tools/test_resume_boundary.py:26-56), a stale CFGSIG (ignored: a fresh
run), a resume from a checkpoint written past a hit or past the last cell, and n = 19
front-only interrupts."""
import os
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

import bs45_quantum_explorer_amd as wz
from golden_data import load

pytestmark = pytest.mark.gpu

ENV0 = {k: v for k, v in os.environ.items() if not k.startswith(('WZ_', 'FH_'))}
SEMANTIC = ('RESULT:', 'candidates_streamed=', 'backtracks_entered=', 'A =', 'B =', 'C =', 'D =', 'VERIFY:',
            'FIRSTHIT:', 'sig =', '[firsthit ckpt]')
CASES = load('resume.json')


def semantic(out):
    return [re.sub(r' elapsed=.*', '', line) for line in out.splitlines() if line.startswith(SEMANTIC)]


def run_cli(n, sig, settings):
    env = {**ENV0, **{k: str(v) for k, v in settings.items()}}
    p = subprocess.run([str(wz.CLIPATH), str(n), *map(str, sig)], env=env, text=True, capture_output=True,
                       timeout=300)
    return p.returncode, p.stdout


@pytest.mark.parametrize('case', CASES, ids=[c['name'].replace(' ', '_') for c in CASES])
def test_checkpoint_and_resume_match_reference(case):
    n, sig = case['n'], case['sig']
    with tempfile.TemporaryDirectory() as td:
        ck = Path(td)
        base = {**case['settings'], 'WZ_FH_CKPT_DIR': ck}
        first = {**base, 'WZ_FH_TEST_STOP_AFTER': case['stop_after']} if case['stop_after'] else base
        rc1, out1 = run_cli(n, sig, first)
        p = ck / 'arm_0.ckpt'
        assert rc1 == case['first']['rc'], out1[-800:]
        assert semantic(out1) == case['first']['lines']
        ck1 = p.read_text() if p.exists() else None
        assert ck1 == case['first']['ckpt'], 'checkpoint bytes after the first run'
        if case['forge'] and ck1 is not None:
            forged = ck1
            for pat, rep in case['forge']:
                forged = re.sub(pat, rep, forged)
            assert forged == case['forged']
            p.write_text(forged)
        rc2, out2 = run_cli(n, sig, base)
        assert rc2 == case['resumed']['rc'], out2[-800:]
        assert semantic(out2) == case['resumed']['lines']
        ck2 = p.read_text() if p.exists() else None
        assert ck2 == case['resumed']['ckpt'], 'checkpoint bytes after the resumed run'


def test_drain_slices_do_not_change_an_arm():
    """The drain's completions go to the device in slices (SIGTERM, hits and the checkpoint clock are
    looked at between slices): one-candidate slices and 4096-candidate slices give the same arm."""
    case = next(c for c in CASES if c['name'].startswith('n19 front stop=37'))
    outs = []
    for smax in (1, 4096):
        with tempfile.TemporaryDirectory() as td:
            st = {**case['settings'], 'WZ_FH_CKPT_DIR': td, 'WZ_FH_TEST_STOP_AFTER': case['stop_after'],
                  'WZFH_DRAIN_SLICE_MAX': smax}
            rc, out = run_cli(case['n'], case['sig'], st)
            outs.append((rc, semantic(out), (Path(td) / 'arm_0.ckpt').read_text()))
    assert outs[0] == outs[1]
    assert outs[0][1] == case['first']['lines'] and outs[0][2] == case['first']['ckpt']
