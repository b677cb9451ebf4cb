"""Frame output in the reference's Three.js JSON schema (FrameWriter.java), so the reference's viewer and file
server (sim_viewer_boa.html, sim_server.py) show runs of this library unchanged.

One file per frame, `frame_NNNNNN.json` in a run folder (FrameWriter.java:62-99):

    {"frame":N,"t":T,"bounds":{"xDim":..,"yDim":..,"zDim":..},
     "segments":[{"id":I,"end1":[x,y,z],"end2":[x,y,z],"r":R,"notADPRatio":..,"cofilinCount":0,
                  "isBarbedEnd":b[,"fATP":..,"fADPPi":..,"fADP":..]}, ...]}

end1 / end2 are derived on the host in double from the pulled pose (coord, uVec, segLength), as the reference
does, so a frame costs one pull of the pose and nothing on the device.  Numbers are formatted as Java's
`String.format(Locale.US, "%.Ng", x)` formats them (every significant digit kept: C's `%#.Ng`).  Java rounds
the decimal expansion half-up where C rounds it to even; the two differ only when a double's exact decimal
expansion ends in a 5 right after the last kept digit.
"""
from __future__ import annotations

import os

import numpy as np

ACTIN_WIDTH_RADIUS = 0.007 / 2.0   # Constants.radius = actinWidth / 2 (Constants.java:39), um


def _g(x: float, digits: int) -> str:
    """Java's %.<digits>g (Locale.US): `digits` significant digits, fixed notation for 1e-4 <= |m| <
    10^digits, scientific otherwise, exponent with at least two digits."""
    s = f"%#.{digits}g" % x
    return s[:-1] if s.endswith(".") else s   # C keeps the point of an integral fixed form ("12346."), Java not


def resolve_output_dir(requested: str) -> str:
    """FrameWriter.resolveOutputDir (FrameWriter.java:45-60): create the folder, or the first free `.NNN`."""
    if not os.path.exists(requested):
        os.makedirs(requested)
        return requested
    for k in range(1, 1000):
        cand = "%s.%03d" % (requested, k)
        if not os.path.exists(cand):
            os.makedirs(cand)
            return cand
    raise RuntimeError(f"FrameWriter: could not allocate an output dir for {requested}")


class FrameWriter:
    """FrameWriter.java: frames of a filament store (and, with turnover, coloured by the nucleotide proxy)."""

    def __init__(self, requested_dir: str, x_dim: float, y_dim: float, z_dim: float):
        self.out_dir = resolve_output_dir(requested_dir)
        self.dims = (x_dim, y_dim, z_dim)
        self.frame_number = 0

    def frames_written(self) -> int:
        return self.frame_number

    def frame_json(self, fs, t: float, turn=None) -> str:
        """The JSON text of one frame (FrameWriter.writeFrame, FrameWriter.java:67-85 and, with a turnover
        store, the aging-coloured overload :95-116: FREE slots skipped, notADPRatio = 1 - f_ADP)."""
        n = fs.n
        c = fs.coord.reshape(3, n).astype(np.float64)
        u = fs.uVec.reshape(3, n).astype(np.float64)
        half = fs.segLength.astype(np.float64) * 0.5
        e1 = c - half * u
        e2 = c + half * u
        barbed = fs.end2NbrSlot < 0
        r = _g(ACTIN_WIDTH_RADIUS, 5)
        parts = ['{"frame":%d' % self.frame_number, ',"t":%s' % _g(t, 6),
                 ',"bounds":{"xDim":%s,"yDim":%s,"zDim":%s}' % tuple(_g(d, 5) for d in self.dims), ',"segments":[']
        segs = []
        if turn is None:
            one = _g(1.0, 3)
            for i in range(n):
                segs.append('{"id":%d,"end1":[%s,%s,%s],"end2":[%s,%s,%s],"r":%s,"notADPRatio":%s,"cofilinCount":0,'
                            '"isBarbedEnd":%s}' % (i, _g(e1[0, i], 5), _g(e1[1, i], 5), _g(e1[2, i], 5),
                                                   _g(e2[0, i], 5), _g(e2[1, i], 5), _g(e2[2, i], 5), r, one,
                                                   "true" if barbed[i] else "false"))
        else:
            C = turn.C
            nf = turn.nucFrac.astype(np.float64)
            for i in range(n):
                if turn.filState[i] < 0:
                    continue
                fATP, fADPPi, fADP = nf[i], nf[C + i], nf[2 * C + i]
                segs.append('{"id":%d,"end1":[%s,%s,%s],"end2":[%s,%s,%s],"r":%s,"notADPRatio":%s,"cofilinCount":0,'
                            '"isBarbedEnd":%s,"fATP":%s,"fADPPi":%s,"fADP":%s}'
                            % (i, _g(e1[0, i], 5), _g(e1[1, i], 5), _g(e1[2, i], 5), _g(e2[0, i], 5), _g(e2[1, i], 5),
                               _g(e2[2, i], 5), r, _g(fATP + fADPPi, 3), "true" if barbed[i] else "false",
                               _g(fATP, 3), _g(fADPPi, 3), _g(fADP, 3)))
        parts.append(",".join(segs))
        parts.append("]}")
        return "".join(parts)

    def write_frame(self, fs, t: float, turn=None) -> str:
        path = os.path.join(self.out_dir, "frame_%06d.json" % self.frame_number)
        with open(path, "w") as f:
            f.write(self.frame_json(fs, t, turn))
        self.frame_number += 1
        return path


def write_network_frame(writer: FrameWriter, net, scene, t: float) -> str:
    """Pull a network's current pose (and turnover state) and write one frame."""
    fs = scene.fs.copy()
    net.pull(fs, shape=True)
    turn = None
    if getattr(net, "turn", None) is not None:
        turn = net.pull_turnover(scene.turn.copy())
    return writer.write_frame(fs, t, turn)


def scan_runs(root: str, max_depth: int = 4):
    """The run folders under root, newest first (sim_server.scan_runs, sim_server.py:44-90): a run folder
    holds frame_000000.json and is not searched further.  Returns [{path, name, frameCount, modified}]."""
    skip = {".git", "node_modules", "__pycache__", "libs", ".idea", ".vscode", "venv", ".venv", "RUN_LOGS",
            "ParameterFiles", "removeMe"}
    runs = []

    def walk(d, depth):
        if os.path.exists(os.path.join(d, "frame_000000.json")):
            frames = [f for f in os.listdir(d) if f.startswith("frame_") and f.endswith(".json")]
            rel = os.path.relpath(d, root).replace(os.sep, "/")
            runs.append({"path": rel, "name": os.path.basename(os.path.abspath(d)), "frameCount": len(frames),
                         "modified": max(os.path.getmtime(os.path.join(d, f)) for f in frames)})
            return
        if depth >= max_depth:
            return
        try:
            entries = sorted(os.listdir(d))
        except OSError:
            return
        for e in entries:
            p = os.path.join(d, e)
            if os.path.isdir(p) and e not in skip and not e.startswith("."):
                walk(p, depth + 1)

    walk(root, 0)
    runs.sort(key=lambda r: r["modified"], reverse=True)
    return runs
