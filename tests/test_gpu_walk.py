"""The chain's partner walk decides each slot from tree-summed partial sums wherever they are provably on
the same side of the target as the walk's sequential partials, and forms the sequential ones otherwise.
Whole operator calls with the sequential walk forced (gdx_col_set_exact_walk) give the same bits:
events, sizes, numbers and draw counters."""
import pytest

pytestmark = pytest.mark.gpu


def _state(model, n, seed):
    from gamedev_amd import synth
    if model == "plain_2d":
        return synth.plain_2d_state(n, seed=seed)
    return synth.fiducial_3d_state(n, seed=seed)


@pytest.mark.parametrize("model", ["plain_2d", "fiducial_3d", "fiducial_k64"])
def test_tree_walk_matches_sequential_walk(gpu, model):
    from gamedev_amd import CollisionOperator, model_info, synth
    st = _state(model, 40000, 13)
    h = None if model == "plain_2d" else 0.05
    outs = []
    for exact in (False, True):
        d = synth.to_device(st, gpu)
        op = CollisionOperator(model, len(st["x"]), st["total_mass"], seed=9)
        op.bind_soa(d["x"], d["y"], d["z"], d["size"], d["numr"])
        op.set_exact_walk(exact)
        op.geometry()
        if h is None:
            # a horizon of several baths: eps^2 / <J2> is about one bath of the controller
            op.rate_pass()
            c = op.cached()
            j2 = c[c[:, 0] > 0, 2].mean()
            h = 8.0*model_info(model)["col_bath_eps"]**2/j2
        ev = []
        for step in range(3):
            print(f"[walk] {model} exact={exact} step {step} h={h:.3e}", flush=True)
            ev.append(op.evolve(h)["events"])
        outs.append((ev, d["size"].cpu().numpy().tobytes(), d["numr"].cpu().numpy().tobytes(), op.rng().tobytes()))
    assert sum(outs[0][0]) > 0
    assert outs[0][0] == outs[1][0]
    assert outs[0][1:] == outs[1][1:]
