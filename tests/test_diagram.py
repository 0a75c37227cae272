"""The FP64 oracle's config.yaml-defaults frame against the summary of the
reference's first-frame artefact (picture/diagram.png); the check and its
thresholds are described in tests/diagram_check.py.  This runs the
reference's as-written path end to end: as-written right-hand side,
secant disk root, tolerance scale floor 0, spp 4, 100-step budget."""
import numpy as np

from kerrtrace_amd import config as K
from oracle import oracle as O
from tests import diagram_check as D


def test_oracle_defaults_frame_matches_reference_artefact():
    cfg, cam = D.scene()
    assert cfg.rhs_mode == K.KT_RHS_REFERENCE and cfg.disk_root_mode == K.KT_DISK_ROOT_SECANT_REF
    assert cfg.tol_floor == 0.0 and cfg.spp == 4 and cfg.max_steps == 100
    hdr, cls, _, _ = O.render(cfg, cam)
    h4 = np.concatenate([hdr, np.ones(hdr.shape[:2] + (1,))], axis=-1).astype(np.float32)
    ldr, _ = O.post(cfg, h4)
    m = D.metrics(ldr)
    print(m)
    assert D.passes(m), m


def test_diagram_check_rejects_a_wrong_scene():
    """The same check fails for the spin's sign flipped (a plausible
    mistake in the host's parameter derivation)."""
    cfg, cam = D.scene()
    cfg.spin = -cfg.spin
    hdr, _, _, _ = O.render(cfg, cam)
    h4 = np.concatenate([hdr, np.ones(hdr.shape[:2] + (1,))], axis=-1).astype(np.float32)
    ldr, _ = O.post(cfg, h4)
    assert not D.passes(D.metrics(ldr))
