"""Checkpoint files for `NewtonAbd`: restartable, bitwise-resumable runs.

The reference keeps its state in Warp arrays and has no file format for it; its state is the BDF1
triple that ``forward_states`` (abd.py:574-581) rolls forward: ``q`` (the Newton iterate,
BDFAffine.nxt), ``q_now`` (the last accepted state) and ``qdot``.  A step of this solver depends
on nothing else that outlives a step except the parameters and the PCG layout (the tile size sets
the order of the PCG's partial sums) and the bodies' contact stiffness (raised by the adaptive rule
of ``Params.k_max``), so a file holding those, the frame number and a fingerprint of the scene
resumes a run bit for bit.  A learned warm start or preconditioner is not part of the
file: a run that uses one resumes bit for bit once the same model is set again.

Format: one ``.npz`` (numpy, no pickled objects), keys

``format``        the string ``colb_amd.abd.checkpoint/1``
``q, q_now, qdot``  ``float64[n_bodies, 12]``
``frame``         ``int64`` frames stepped so far
``params``        JSON of :class:`colb_amd.params.Params` (tuples as lists)
``pcg_layout``    ``int32[2]``: bodies per tile, one launch per iteration (0/1)
``stiffness``     ``float64[n_bodies]`` contact stiffness of each body (absent in older files:
                  every body at ``params.k_contact``)
``scene_sha256``  SHA-256 over the scene's geometry, topology, masses, pins and scene ids
"""
import hashlib
import json

import numpy as np

from .params import Params

FORMAT = "colb_amd.abd.checkpoint/1"


def scene_fingerprint(scene):
    """SHA-256 (hex) of everything that defines the bodies of a scene except their state."""
    h = hashlib.sha256()
    for a in (scene.rest_x, scene.vert_body, scene.edges, scene.tris, scene.body_scene, scene.body_pinned,
              scene.mass4):
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode())
        h.update(np.asarray(a.shape, dtype=np.int64).tobytes())
        h.update(a.tobytes())
    return h.hexdigest()


def save(solver, path):
    """Write the state of ``solver`` (a NewtonAbd) to ``path`` (an .npz file)."""
    q, qn, qd = solver.get_state()
    bpt, multi = solver.pcg_layout
    with open(path, "wb") as f:
        np.savez(f, format=np.array(FORMAT), q=q, q_now=qn, qdot=qd, frame=np.int64(solver.frame),
                 params=np.array(json.dumps(solver.params.as_dict())),
                 pcg_layout=np.array([bpt, int(bool(multi))], dtype=np.int32),
                 stiffness=np.asarray(solver.stiffness(), dtype=np.float64),
                 scene_sha256=np.array(scene_fingerprint(solver.scene)))


def read(path):
    """The checkpoint at ``path`` as a dict (arrays, ``frame`` int, ``params`` Params)."""
    with np.load(path, allow_pickle=False) as z:
        if "format" not in z.files or str(z["format"]) != FORMAT:
            raise ValueError(f"{path}: not a {FORMAT} file")
        d = {k: z[k] for k in z.files}
    raw = json.loads(str(d["params"]))
    for k, v in raw.items():
        if isinstance(v, list):
            raw[k] = tuple(v)
    return {"q": d["q"], "q_now": d["q_now"], "qdot": d["qdot"], "frame": int(d["frame"]),
            "params": Params(**raw), "pcg_layout": (int(d["pcg_layout"][0]), bool(d["pcg_layout"][1])),
            "stiffness": d.get("stiffness"), "scene_sha256": str(d["scene_sha256"])}


def load(solver, path):
    """Restore ``solver`` (built on the same scene) from ``path``: parameters, PCG layout, state
    and frame.  Raises ValueError when the file was written for another scene."""
    d = read(path)
    nb = solver.n_bodies
    if d["scene_sha256"] != scene_fingerprint(solver.scene):
        raise ValueError(f"{path}: checkpoint of another scene (fingerprint differs)")
    for k in ("q", "q_now", "qdot"):
        if d[k].shape != (nb, 12):
            raise ValueError(f"{path}: {k} has shape {d[k].shape}, expected {(nb, 12)}")
    solver.set_params(d["params"])
    solver.set_pcg_layout(*d["pcg_layout"])
    solver.set_state(d["q"], d["q_now"], d["qdot"])
    if d["stiffness"] is not None:
        if d["stiffness"].shape != (nb,):
            raise ValueError(f"{path}: stiffness has shape {d['stiffness'].shape}, expected {(nb,)}")
        solver.set_stiffness(d["stiffness"])
    solver.frame = d["frame"]
    solver.compute_V()
    return d
