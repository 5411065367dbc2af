"""Writes tests/golden/hp_vectors.json: high-precision (mpmath, 40 digits)
values that pin the oracle's closed forms without trusting its arithmetic.

  geodetic_to_ecef  — WGS84 closed form (geodesy.rs:38-55) at 12 points
  lagrange8         — degree-7 Lagrange pos/vel/accel (orbit.rs:498-596) of a
                      synthetic orbit at 10 times, with the reference's window
                      rule, evaluated in exact rational-like 40-digit arithmetic
                      (accel = exact derivative of the velocity polynomial)
  utm               — Krüger n^6 forward TM at 6 points (Karney 2011 series)
  meridian_arc      — northing on a central meridian, k0·M(φ), with the
                      meridian arc M(φ) = a(1-e²)∫_0^φ (1-e² sin²t)^(-3/2) dt
                      integrated by mpmath quadrature (no series involved)

Run: python tests/golden/make_vectors.py   (needs mpmath; numpy for the orbit)
"""
import json
import pathlib
import sys

import mpmath as mp

mp.mp.dps = 40
HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

A = mp.mpf(6378137)
F = 1 / mp.mpf("298.257223563")
E2 = 2 * F - F * F


def ecef(lat, lon, h):
    la = mp.mpf(lat) * mp.pi / 180
    lo = mp.mpf(lon) * mp.pi / 180
    n = A / mp.sqrt(1 - E2 * mp.sin(la) ** 2)
    return [(n + h) * mp.cos(la) * mp.cos(lo), (n + h) * mp.cos(la) * mp.sin(lo), (n * (1 - E2) + h) * mp.sin(la)]


def lagrange(ts, P, V, t):
    n = len(ts)
    # centre / window rule of orbit.rs:522-535
    c = max(i for i in range(n) if ts[i] <= t)
    if c < 4:
        s = 0
    elif c + 4 >= n:
        s = n - 8
    else:
        s = c - 3
    tw = [mp.mpf(ts[s + j]) / 10**6 for j in range(8)]
    x = mp.mpf(t) / 10**6
    pos = [mp.mpf(0)] * 3
    vel = [mp.mpf(0)] * 3
    acc = [mp.mpf(0)] * 3
    for k in range(8):
        L = mp.mpf(1)
        for j in range(8):
            if j != k:
                L *= (x - tw[j]) / (tw[k] - tw[j])
        dL = mp.mpf(0)
        for m in range(8):
            if m == k:
                continue
            term = 1 / (tw[k] - tw[m])
            for j in range(8):
                if j != k and j != m:
                    term *= (x - tw[j]) / (tw[k] - tw[j])
            dL += term
        for c3 in range(3):
            pos[c3] += L * mp.mpf(P[s + k][c3])
            vel[c3] += L * mp.mpf(V[s + k][c3])
            acc[c3] += dL * mp.mpf(V[s + k][c3])
    return pos, vel, acc


def utm_fwd(zone, lon, lat):
    n = F / (2 - F)
    A_ = A / (1 + n) * (1 + n**2 / 4 + n**4 / 64 + n**6 / 256)
    al = [n / 2 - 2 * n**2 / 3 + 5 * n**3 / 16 + 41 * n**4 / 180 - 127 * n**5 / 288 + 7891 * n**6 / 37800,
          13 * n**2 / 48 - 3 * n**3 / 5 + 557 * n**4 / 1440 + 281 * n**5 / 630 - 1983433 * n**6 / 1935360,
          61 * n**3 / 240 - 103 * n**4 / 140 + 15061 * n**5 / 26880 + 167603 * n**6 / 181440,
          49561 * n**4 / 161280 - 179 * n**5 / 168 + 6601661 * n**6 / 7257600,
          34729 * n**5 / 80640 - 3418889 * n**6 / 1995840,
          212378941 * n**6 / 319334400]
    e = mp.sqrt(E2)
    phi = mp.mpf(lat) * mp.pi / 180
    lam = (mp.mpf(lon) - (zone * 6 - 183)) * mp.pi / 180
    t = mp.sinh(mp.atanh(mp.sin(phi)) - e * mp.atanh(e * mp.sin(phi)))
    xi_ = mp.atan2(t, mp.cos(lam))
    eta_ = mp.atanh(mp.sin(lam) / mp.sqrt(1 + t * t))
    xi = xi_ + sum(al[j] * mp.sin(2 * (j + 1) * xi_) * mp.cosh(2 * (j + 1) * eta_) for j in range(6))
    eta = eta_ + sum(al[j] * mp.cos(2 * (j + 1) * xi_) * mp.sinh(2 * (j + 1) * eta_) for j in range(6))
    k0 = mp.mpf("0.9996")
    return 500000 + k0 * A_ * eta, k0 * A_ * xi


def meridian_northing(lat):
    phi = mp.mpf(lat) * mp.pi / 180
    m = A * (1 - E2) * mp.quad(lambda t: (1 - E2 * mp.sin(t) ** 2) ** mp.mpf(-1.5), [0, phi])
    k0 = mp.mpf("0.9996")
    return k0 * m + (10_000_000 if lat < 0 else 0)


def main():
    from sardine_amd import synthetic

    out = {}
    pts = [(0, 0, 0), (90, 0, 0), (51.5, 0, 0), (50.5, 10, 300), (0, 45, 693000), (45.3, 9.7, 1234.5),
           (-33.9, 151.2, 20), (89.9, -179.9, 5), (-89.5, 120, -50), (12.34567, -98.7654, 2500),
           (47.1, 11.6, 700000), (60, 180, 0)]
    out["geodetic_to_ecef"] = [{"lat": a, "lon": b, "h": c, "xyz": [str(v) for v in ecef(a, b, c)]} for a, b, c in pts]
    sc = synthetic.scene_small_plain(0, lines=64, samples=64, out_rows=8, out_cols=8, dem_side=8)
    ts = [int(t - sc.sv_t_us[0]) for t in sc.sv_t_us]
    P = [[float(x) for x in p] for p in sc.sv_pos]
    V = [[float(x) for x in v] for v in sc.sv_vel]
    times = [0, 5_000_000, 25_123_456, 39_999_999, 40_000_000, 78_950_001, 80_000_000, 123_456_789,
             155_000_000, 160_000_000]
    lg = []
    for t in times:
        p, v, a = lagrange(ts, P, V, t)
        lg.append({"t_rel_us": t, "pos": [str(x) for x in p], "vel": [str(x) for x in v], "acc": [str(x) for x in a]})
    out["lagrange8"] = {"sv_t_rel_us": ts, "pos": P, "vel": V, "cases": lg}
    ut = [(32, 11.5, 48.1), (32, 9.0, 45.0), (33, 15.0, 0.5), (31, 0.5, 60.0), (32, 12.9, 47.9), (1, -177.0, 10.0)]
    out["utm_north"] = [{"zone": z, "lon": lo, "lat": la, "xy": [str(v) for v in utm_fwd(z, lo, la)]}
                        for z, lo, la in ut]
    lats = [-80.0, -45.5, -10.0, -0.25, 0.5, 30.0, 48.1374, 60.0, 72.5, 84.0]
    out["meridian_arc"] = [{"lat": la, "northing": str(meridian_northing(la))} for la in lats]
    (HERE / "hp_vectors.json").write_text(json.dumps(out, indent=1))
    print("wrote", HERE / "hp_vectors.json")


if __name__ == "__main__":
    main()
