"""Tuning knobs of the batched path: the autotuner's file is read by new contexts,
and the work distribution changes no result."""
import json
import os

import numpy as np
import pytest

import hbtx
import synth

pytestmark = pytest.mark.gpu


def test_tune_file_is_read(tmp_path):
    f = tmp_path / "tune.json"
    f.write_text(json.dumps({"direct_cutoff": 777, "walk_chunks": 4, "cutoff_timings": []}))
    old = os.environ.get("HBTX_TUNE_FILE")
    os.environ["HBTX_TUNE_FILE"] = str(f)
    try:
        c = hbtx.Context(0)
        assert c.tuning() == dict(direct_cutoff=777, walk_chunks=4)
        c.set_tuning(100, 2)
        assert c.tuning() == dict(direct_cutoff=100, walk_chunks=2)
        with pytest.raises(hbtx.HbtxError):
            c.set_tuning(-1, 2)
        with pytest.raises(hbtx.HbtxError):
            c.set_tuning(10, 0)
        c.close()
    finally:
        if old is None:
            del os.environ["HBTX_TUNE_FILE"]
        else:
            os.environ["HBTX_TUNE_FILE"] = old


def test_walk_chunking_changes_no_result(ctx):
    pop = synth.population(seed=41, nsub=300, nmax=40000)
    prm = hbtx.workload_params(direct_cutoff=-1)
    saved = ctx.tuning()
    outs = []
    try:
        for ch in (1, 8, 13):
            ctx.set_tuning(saved["direct_cutoff"], ch)
            outs.append(ctx.unbind(pop, prm))
    finally:
        ctx.set_tuning(saved["direct_cutoff"], saved["walk_chunks"])
    for o in outs[1:]:
        assert np.array_equal(o["order"], outs[0]["order"])
        assert np.array_equal(o["energy"], outs[0]["energy"])
        assert np.array_equal(o["result"], outs[0]["result"])


def test_autotuner_runs_end_to_end(tmp_path):
    """The autotuner on a small budget (2e5 particles per population): every
    equal-size population, the power-law refinement and the chunk sweep run, and
    the file it writes holds one of its candidates.  Its 2048-source populations
    have about 2250 sources per subhalo, segments just over the LDS sort size."""
    import subprocess
    import sys
    out = tmp_path / "tune.json"
    script = os.path.join(os.path.dirname(hbtx.LIB_PATH), "..", "autotune.py")
    print("autotune --particles 2e5", flush=True)
    r = subprocess.run([sys.executable, script, "--particles", "2e5", "--out", str(out)], capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(out.read_text())
    cands = {row["direct_cutoff"] for row in d["mixture_timings"]}
    assert d["direct_cutoff"] in cands and d["walk_chunks"] in (1, 4, 8, 16)
    assert len(d["cutoff_timings"]) == 8
