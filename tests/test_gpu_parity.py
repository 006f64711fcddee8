"""GPU substep vs the CPU oracle (f32 restatement of Simulation::step) and the
FP64 oracle, per particle, after 1 step and after the scene's N steps."""
import numpy as np
import pytest

import scenes as S


def _run(scene, backend, steps):
    sim = scene.build(backend)
    out = []
    done = 0
    for s in steps:
        sim.step_n(s - done)
        done = s
        if backend == "oracle64":
            out.append(S.f64_state(sim.particles_f64()))
        else:
            out.append(S.state_matrix(sim.particles()))
    return out, sim


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(S.SCENES))
def test_scene_parity(name, oracle, capsys):
    scene = S.SCENES[name]
    steps = [1, scene.steps]
    g, gsim = _run(scene, "gpu", steps)
    o32, osim = _run(scene, "oracle32", steps)
    o64, _ = _run(scene, "oracle64", steps)
    for k, s in enumerate(steps):
        e_g32 = S.rel_errors(g[k], o32[k])
        e_g64 = S.rel_errors(g[k], o64[k])
        e_3264 = S.rel_errors(o32[k], o64[k])
        with capsys.disabled():
            print(f"\n[{name} after {s} steps] n={len(g[k]['x'])}")
            for f in ("x", "v", "C", "F", "Jp", "q", "lvs", "density"):
                print(f"  {f:8s} gpu-vs-o32 {e_g32[f]:.3e}  gpu-vs-o64 {e_g64[f]:.3e}  o32-vs-o64 {e_3264[f]:.3e}")
    assert gsim.stats().total_substeps == osim.stats().total_substeps
