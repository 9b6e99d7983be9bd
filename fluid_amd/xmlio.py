"""The reference's particle XML format (the ``-p`` input of
src/application.cpp:302-344):

    <particles><density>rho0</density><ps>
      <particle><pos>x y z</pos><v>vx vy vz</v></particle> ...
    </ps></particles>

Positions and velocities are parsed as double like ``stov``
(src/application.cpp:293-300); the density as float like ``stof`` (:331).
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np


def load_particles(path: str):
    """Returns (rest_density, positions (n,3) float64, velocities (n,3) float64)."""
    root = ET.parse(path).getroot()
    if root.tag != "particles":
        raise ValueError(f"{path}: not a particles file (root <{root.tag}>)")
    dens = root.find("density")
    ps = root.find("ps")
    if dens is None or ps is None:
        raise ValueError(f"{path}: <density> and <ps> are required")
    rho0 = float(np.float32(float(dens.text.strip())))
    pos, vel = [], []
    for p in ps.findall("particle"):
        pos.append([float(t) for t in p.find("pos").text.split()[:3]])
        vel.append([float(t) for t in p.find("v").text.split()[:3]])
    return rho0, np.asarray(pos, np.float64).reshape(-1, 3), np.asarray(vel, np.float64).reshape(-1, 3)


def save_particles(path: str, rest_density: float, x: np.ndarray, v: np.ndarray) -> None:
    """Writes the same format (the reference has no writer)."""
    with open(path, "w") as f:
        f.write("<particles>\n  <density>%r</density>\n  <ps>\n" % float(rest_density))
        for p, q in zip(np.asarray(x, np.float64).tolist(), np.asarray(v, np.float64).tolist()):
            f.write("    <particle>\n      <pos>%r %r %r</pos>\n      <v>%r %r %r</v>\n    </particle>\n"
                    % (p[0], p[1], p[2], q[0], q[1], q[2]))
        f.write("  </ps>\n</particles>\n")
