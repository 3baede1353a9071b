#!/usr/bin/env python3
"""Writes c2p_conformal.json: conserved states Q and the exact primitive
solution of the conformal (p = e/3) con2prim problem, from the closed form
M v^2 - 4 Q_t v + 3 M = 0  ->  v = 3M / (2 Q_t + sqrt(4 Q_t^2 - 3 M^2)),
evaluated with Python's decimal module at 50 digits (independent of numpy and
of the oracle).  The reference's transformPV (src/rmn.cpp:28-139) solves the
same equation iteratively and stops at |dv| < 1e-5.

Also records inputs with |Q_vec| > Q_t: the reference first rescales Q_vec to
0.9999 Q_t (src/rmn.cpp:65-70), so the exact answer is that of the rescaled
state."""
import json
import os
import random
from decimal import Decimal, getcontext

getcontext().prec = 50
random.seed(20240611)


def exact(Qt, Qx, Qy, Qz, nb):
    Qt, Qx, Qy, Qz, nb = (Decimal(repr(a)) for a in (Qt, Qx, Qy, Qz, nb))
    M = (Qx * Qx + Qy * Qy + Qz * Qz).sqrt()
    if Qt <= 0:
        return None
    if M == 0:
        return dict(e=Qt, v=Decimal(0), vx=Decimal(0), vy=Decimal(0), vz=Decimal(0), nb=nb)
    if M > Qt:
        f = Decimal("0.9999") * Qt / M
        Qx, Qy, Qz = Qx * f, Qy * f, Qz * f
        M = Qt * Decimal("0.9999")
    v = 3 * M / (2 * Qt + (4 * Qt * Qt - 3 * M * M).sqrt())
    e = Qt - M * v
    return dict(e=e, v=v, vx=v * Qx / M, vy=v * Qy / M, vz=v * Qz / M,
                nb=nb * (1 - v * v).sqrt())


cases = []
for i in range(400):
    Qt = 10 ** random.uniform(-5, 3)
    frac = [0.0, 1e-8, 0.3, 0.9, 0.99, 0.999, 0.99999, 1.2][i % 8] if i < 80 else random.random()
    M = frac * Qt
    th = random.uniform(0, 3.14159)
    ph = random.uniform(0, 6.28318)
    import math
    Qx, Qy, Qz = (M * math.sin(th) * math.cos(ph), M * math.sin(th) * math.sin(ph),
                  M * math.cos(th))
    nb = random.choice([0.0, random.uniform(0, 2)])
    sol = exact(Qt, Qx, Qy, Qz, nb)
    cases.append(dict(Q=[Qt, Qx, Qy, Qz, nb, 0.0, 0.0],
                      exact={k: str(v) for k, v in sol.items()}))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2p_conformal.json")
with open(out, "w") as f:
    json.dump(dict(source="closed-form conformal root, decimal 50 digits", cases=cases), f,
              indent=0)
print("wrote", out, len(cases))
