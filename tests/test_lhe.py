"""Les Houches output (madflow_amd/lhe.py): the written file reads back with the same
events, momenta and process information; the layout follows the Les Houches Accord."""

import gzip

import numpy as np
import pytest

from madflow_amd.lhe import write_lhe, read_lhe, PDG
from tests.helpers import phase_space_points


@pytest.mark.parametrize("name", ["gg_ttx", "gg_ttxgg"])
def test_lhe_round_trip(tmp_path, name):
    ps, _ = phase_space_points(name, 50, seed=3)
    path = tmp_path / ("events_%s.lhe.gz" % name)
    n = write_lhe(path, name, ps, sigma_pb=12.5, err_pb=0.25, sqrts=ps[0, 0, 0] * 2)
    assert n == 50
    with gzip.open(path, "rt") as f:
        text = f.read()
    assert text.startswith('<LesHouchesEvents version="3.0">') and text.rstrip().endswith("</LesHouchesEvents>")
    assert text.count("<event>") == 50 and text.count("</event>") == 50
    init, events = read_lhe(path)
    beam, proc = init
    assert beam[:2] == [21.0, 21.0] and beam[2] == beam[3] == ps[0, 0, 0]
    assert beam[8] == 3.0                                   # IDWTUP: unweighted events
    assert proc == [12.5, 0.25, 12.5, 1.0]
    assert len(events) == 50
    for ev, (head, parts) in zip(ps, events):
        assert head[0] == len(PDG[name]) and head[2] == 12.5
        assert [p[0] for p in parts] == PDG[name]
        assert [p[1] for p in parts] == [-1, -1] + [1] * (len(PDG[name]) - 2)
        got = np.array([[p[5], p[2], p[3], p[4]] for p in parts])
        np.testing.assert_allclose(got, ev, rtol=1e-11, atol=1e-9)
        masses = [p[6] for p in parts]
        assert abs(masses[2] - 173.0) < 1e-6 and abs(masses[3] - 173.0) < 1e-6


def test_lhe_rejects_wrong_shape(tmp_path):
    with pytest.raises(ValueError):
        write_lhe(tmp_path / "x.lhe", "gg_ttxg", np.zeros((3, 4, 4)), 1.0, 0.1, 1000.0)


def test_lhe_overweight_events_keep_their_weight(tmp_path):
    """Events kept with w > wmax (hit-or-miss keeps them always) carry sigma * w / wmax, the
    others sigma; the file then declares weighted events (IDWTUP = 4)."""
    ps, _ = phase_space_points("gg_ttxg", 6, seed=4)
    w = np.array([0.5, 2.0, 1.0, 3.5, 0.1, 1.0])
    path = tmp_path / "ow.lhe"
    write_lhe(path, "gg_ttxg", ps, sigma_pb=10.0, err_pb=0.1, sqrts=2000.0, weights=w, wmax=1.0)
    init, events = read_lhe(path)
    assert init[0][8] == 4.0
    np.testing.assert_allclose([ev[0][2] for ev in events], [10.0, 20.0, 10.0, 35.0, 10.0, 10.0], rtol=1e-12)
    with open(path) as f:
        assert 'overweight_events="2"' in f.read()
    # without over-weight events the file stays unweighted
    write_lhe(path, "gg_ttxg", ps, sigma_pb=10.0, err_pb=0.1, sqrts=2000.0, weights=np.minimum(w, 1.0), wmax=1.0)
    init, events = read_lhe(path)
    assert init[0][8] == 3.0 and all(ev[0][2] == 10.0 for ev in events)
    with pytest.raises(ValueError):
        write_lhe(path, "gg_ttxg", ps, 10.0, 0.1, 2000.0, weights=w)


def test_every_generated_process_has_pdg_codes():
    """write_lhe knows the external legs of every process the generator emits: one PDG code
    per leg, gluons 21, tops +-6, light quarks +-2."""
    from madflow_amd.lhe import PDG
    from madflow_amd.procgen.emit import descriptions
    code = {"gluon": 21, "top": 6, "atop": -6, "quark": 2, "antiquark": -2, "quark_out": 2, "antiquark_out": -2}
    for d in descriptions():
        assert PDG[d["name"]] == [code[k] for k in d["kinds"]], d["name"]


def _writer_batches(writer, name, nev, seed):
    """Two batches of weighted events handed to lhe_parser as the integrand does."""
    ps, _ = phase_space_points(name, nev, seed=seed)
    w = np.random.default_rng(seed).random(nev) ** 3
    half = nev // 2
    assert writer.lhe_parser(ps[:half], w[:half]) == 0.0
    assert writer.lhe_parser(ps[half:], w[half:]) == 0.0
    return ps, w


def test_lhe_writer_weighted_events(tmp_path):
    """LheWriter(no_unweight=True): the weighted events of every lhe_parser call reach
    Events/<run>/weighted_events.lhe.gz in order, each with its own weight (IDWTUP 4), and
    the stored result heads the file; dump_result writes (sigma, err)."""
    from madflow_amd.lhe import LheWriter
    with LheWriter(tmp_path, run="run_07", no_unweight=True, process="gg_ttxg", sqrts=2000.0) as wr:
        ps, w = _writer_batches(wr, "gg_ttxg", 40, seed=5)
        wr.store_result((3.25, 0.125))
    assert wr.cross == 3.25 and wr.err == 0.125 and wr.stats is None
    path = tmp_path / "Events" / "run_07" / "weighted_events.lhe.gz"
    assert path == wr.lhe_path and path.exists() and not wr.unweighted_path.exists()
    init, events = read_lhe(path)
    assert init[0][8] == 4.0 and init[1] == [3.25, 0.125, 3.25, 1.0]
    assert len(events) == 40
    np.testing.assert_allclose([ev[0][2] for ev in events], w, rtol=1e-10)
    got = np.array([[[p[5], p[2], p[3], p[4]] for p in parts] for _, parts in events])
    np.testing.assert_allclose(got, ps, rtol=1e-11, atol=1e-9)
    wr.dump_result(tmp_path / "result.txt")
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "result.txt"), [3.25, 0.125])
    with pytest.raises(ValueError):
        LheWriter(tmp_path, process="pp_ttx", sqrts=2000.0)
    with pytest.raises(ValueError):
        LheWriter(tmp_path, process="gg_ttx", no_unweight=True)
    with pytest.raises(ValueError):
        with LheWriter(tmp_path, process="gg_ttx", sqrts=1000.0, no_unweight=True) as wr:
            wr.lhe_parser(np.zeros((3, 4, 4)), np.zeros(2))


@pytest.mark.gpu
@pytest.mark.parametrize("target", [0, 3000, 9000])
def test_lhe_writer_unweights_on_the_device(tmp_path, dev, target):
    """On leaving the block the weighted events are unweighted on the device: the kept
    count is within 5 sigma of the hit-or-miss expectation sum min(1, w / wmax), where wmax
    is the largest weight or (event_target) the weight whose expectation is the target;
    every kept event is one of the weighted events and carries the cross section."""
    from madflow_amd.lhe import LheWriter
    nev = 20000
    print("case LheWriter unweighting target=%d" % target, flush=True)
    with LheWriter(tmp_path, event_target=target, process="gg_ttx", sqrts=1000.0, seed=11) as wr:
        ps, w = _writer_batches(wr, "gg_ttx", nev, seed=9)
        wr.store_result((7.5, 0.5))
    kept, total = wr.stats
    assert total == nev
    init, events = read_lhe(wr.unweighted_path)
    assert len(events) == kept and init[1][0] == 7.5
    if target:
        at_max = float((w / w.max()).sum())
        # 3000 is below what the largest weight keeps (about 5000): wmax rises above it and
        # every event is unweighted; 9000 lowers wmax and leaves over-weight events
        assert init[0][8] == (4.0 if target > at_max else 3.0)
    else:
        expect = float((w / w.max()).sum())
        assert init[0][8] == 3.0 and all(ev[0][2] == 7.5 for ev in events)
        target = expect
    assert abs(kept - target) < 5 * np.sqrt(target)
    assert all(ev[0][2] >= 7.5 for ev in events)
    order = np.argsort(ps[:, 2, 3])                # the top's pz (its energy is fixed for 2 -> 2)
    pz_sorted = ps[order, 2, 3]
    for _, parts in events:
        mom = np.array([[p[5], p[2], p[3], p[4]] for p in parts])
        i = int(np.searchsorted(pz_sorted, mom[2, 3]))
        near = order[max(0, i - 2):i + 2]
        assert any(np.allclose(ps[j], mom, rtol=1e-11, atol=1e-9) for j in near)


def test_wmax_for_target_hand_solved():
    """w = (1, 2, 3, 4): the largest weight keeps 1/4 + 2/4 + 3/4 + 1 = 2.5 events on
    average.  A target of 2 needs w_max above 4, where the count is 10 / w_max: w_max = 5.
    A target of 3 puts w_max in [3, 4], where it is 1/w + 2/w + 1 + 1 = 3 at w_max = 3; a
    target of 3.5 gives 3/w + 2 = 3.5 in [2, 3], w_max = 2.  No target (or one of at least the
    number of weights) keeps the largest weight."""
    from madflow_amd.lhe import wmax_for_target
    w = np.array([1.0, 2.0, 3.0, 4.0])
    assert wmax_for_target(w) == 4.0 and wmax_for_target(w, 4) == 4.0 and wmax_for_target(w, 9) == 4.0
    assert wmax_for_target(w, 2) == pytest.approx(5.0, rel=1e-14)
    assert wmax_for_target(w, 2.5) == pytest.approx(4.0, rel=1e-14)
    assert wmax_for_target(w, 3) == pytest.approx(3.0, rel=1e-12)
    assert wmax_for_target(w, 3.5) == pytest.approx(2.0, rel=1e-12)
