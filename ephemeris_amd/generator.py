"""Seeded synthetic systems written in Ephemeris' text input formats.

Every file follows the formats the reference loader parses
(src/modules/mass_loader.cpp:67-294 for the config file, :324-561 for the
Initial and Extra files, src/physics/geopotential.cpp:8-127 and
src/physics/ring.cpp:25-108 for the model files; described in
initials/README.md).  Numbers are printed with ``repr`` (shortest round-trip
form), so every loader that uses ``strtod`` reads back the exact doubles.

Workloads (sizes from the benchmark definition):

* ``make_small``    -- 25 bodies: star (J2=2e-7) + 8 planets + 16 moons,
  Newtonian+1PN+J2 only; integrated 1 day at dt=300 s.
* ``make_full``     -- 504 bodies: star, 8 planets with 77 moons, 5 dwarf
  planets with 8 moons, 343 main-belt asteroids (one is a dwarf), 31 TNOs
  (four are dwarfs), 36 soft-mass Kuiper ring points; geopotential J3-J6 on
  star and planets, constant-time-lag tides on planets and major moons, one
  planetary ring, luminosity, mass loss, receptance and J2 rate.
* ``make_ensemble`` -- clones of the full system truncated to the N most
  massive bodies (N ~ U(64,504)) with 1e-9 relative Gaussian perturbations.
* ``make_field9`` / ``make_particles`` -- the star and 8 planets of the full
  system (Newtonian+1PN, solar J2) and asteroid-like massless test particles.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field, replace
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

G = 6.67430e-11
AU = 1.495978707e11
DAY = 86400.0
YEAR = 365.25 * DAY
M_SUN = 1.989e30

# column order of the Initial file (initials/README.md, "Initial File")
INITIAL_HEADER = ("#Name\tID\tGM\tRadius(km)\tinertia\tk2\tk2r\ttd\tJ2\tC21\tC22\tS21\tS22"
                  "\tTidalDelayFactor\tGeopotential\tRing\tx\ty\tz\tvx\tvy\tvz"
                  "\talpha\tdelta\tw\tperiod(h)\txx\txy\txz\tzx\tzy\tzz\twx\twy\twz")


@dataclass
class Body:
    name: str
    sid: str
    GM: float
    R_km: float
    inertia: float = 0.4
    k2: float = 0.0
    k2r: float = 0.0
    td: float = 0.0
    J2: float = 0.0
    C21: float = 0.0
    C22: float = 0.0
    S21: float = 0.0
    S22: float = 0.0
    tdf: float = 1.0
    gp: float = 0.0
    ring: float = 0.0
    r: np.ndarray = field(default_factory=lambda: np.zeros(3))
    v: np.ndarray = field(default_factory=lambda: np.zeros(3))
    x_axis: np.ndarray = field(default_factory=lambda: np.array([1.0, 0.0, 0.0]))
    z_axis: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 1.0]))
    w: np.ndarray = field(default_factory=lambda: np.zeros(3))
    extra: dict = field(default_factory=dict)
    gp_terms: Optional[list] = None      # [(n, m, C, S)]
    rings: Optional[list] = None         # [(name, rin_km, rout_km, thick_km, GM)]


@dataclass
class SystemSpec:
    bodies: List[Body]
    delta_t: float = 300.0
    tdb: int = 0

    @property
    def n(self):
        return len(self.bodies)


def _fmt(x: float) -> str:
    return repr(float(x))


def write_system(spec: SystemSpec, out_dir, stem: str = "System") -> Path:
    """Write Config/Initial/Extra/Geopotentials/Rings files; return the
    config path."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    init_name, extra_name = f"{stem}_Initial.txt", f"{stem}_Extra.txt"
    lines = [INITIAL_HEADER]
    extras = ["#sid\tparam\tvalue"]
    has_gp = has_ring = False
    for b in spec.bodies:
        assert "\t" not in b.name and len(b.sid.encode()) < 8 and " " not in b.sid
        vals = [b.GM, b.R_km, b.inertia, b.k2, b.k2r, b.td, b.J2, b.C21, b.C22, b.S21,
                b.S22, b.tdf, b.gp, b.ring, *b.r, *b.v, 0.0, 90.0, 0.0, 24.0,
                *b.x_axis, *b.z_axis, *b.w]
        lines.append(b.name + "\t" + b.sid + "\t" + "\t".join(_fmt(v) for v in vals))
        for k, v in b.extra.items():
            extras.append(f"{b.sid}\t{k}\t{_fmt(v)}")
        if b.gp:
            has_gp = True
            gdir = out / "Geopotentials"
            gdir.mkdir(exist_ok=True)
            nz = max(t[0] for t in b.gp_terms)
            nt = max([t[0] for t in b.gp_terms if t[1] > 0] or [0])
            gl = [f"{nz}\t{nt}"]
            for n, m, c, s in b.gp_terms:
                gl.append(f"{n}\t{m}\t{_fmt(c)}\t{_fmt(s)}" if m else f"{n}\t0\t{_fmt(c)}")
            (gdir / f"{b.sid}.txt").write_text("\n".join(gl) + "\n")
        if b.ring:
            has_ring = True
            rdir = out / "Rings"
            rdir.mkdir(exist_ok=True)
            rl = ["#name\tRin(km)\tRout\tThick\tGM"]
            for nm, rin, rout, th, gm in b.rings:
                rl.append(f"{nm}\t{_fmt(rin)}\t{_fmt(rout)}\t{_fmt(th)}\t{_fmt(gm)}")
            (rdir / f"{b.sid}.txt").write_text("\n".join(rl) + "\n")
    (out / init_name).write_text("\n".join(lines) + "\n")
    cfg = [f"Initial\t{init_name}"]
    if len(extras) > 1:
        (out / extra_name).write_text("\n".join(extras) + "\n")
        cfg.append(f"Extra\t{extra_name}")
    if has_gp:
        cfg.append("Geopotentials\tGeopotentials")
    if has_ring:
        cfg.append("Rings\tRings")
    cfg += [f"TDB\t{spec.tdb}", "Integrator\tGPU_RK12", f"Delta_t\t{int(spec.delta_t)}",
            f"Cadence\t{int(spec.delta_t)}"]
    path = out / f"{stem}_Config.txt"
    path.write_text("\n".join(cfg) + "\n")
    return path


# --------------------------------------------------------------------------
# orbital mechanics helpers (float64 numpy)
# --------------------------------------------------------------------------
def _rot(i, Om, om):
    cO, sO, ci, si, co, so = math.cos(Om), math.sin(Om), math.cos(i), math.sin(i), math.cos(om), math.sin(om)
    return np.array([[cO * co - sO * so * ci, -cO * so - sO * co * ci, sO * si],
                     [sO * co + cO * so * ci, -sO * so + cO * co * ci, -cO * si],
                     [so * si, co * si, ci]])


def kepler_state(mu, a, e, i, Om, om, M):
    E = M if e < 0.8 else math.pi
    for _ in range(60):
        dE = (E - e * math.sin(E) - M) / (1 - e * math.cos(E))
        E -= dE
        if abs(dE) < 1e-15:
            break
    cE, sE = math.cos(E), math.sin(E)
    r = a * (1 - e * cE)
    x, y = a * (cE - e), a * math.sqrt(1 - e * e) * sE
    fac = math.sqrt(mu * a) / r
    vx, vy = -fac * sE, fac * math.sqrt(1 - e * e) * cE
    Rm = _rot(i, Om, om)
    return Rm @ np.array([x, y, 0.0]), Rm @ np.array([vx, vy, 0.0]), Rm[:, 2]


def _spin_frame(rng, axis, period_h, tilt_max):
    """Body frame with z tilted from ``axis`` by up to ``tilt_max`` rad."""
    axis = axis / np.linalg.norm(axis)
    tmp = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(axis, tmp); u /= np.linalg.norm(u)
    v = np.cross(axis, u)
    t, p = rng.uniform(0, tilt_max), rng.uniform(0, 2 * math.pi)
    z = math.cos(t) * axis + math.sin(t) * (math.cos(p) * u + math.sin(p) * v)
    z /= np.linalg.norm(z)
    tmp = np.array([1.0, 0.0, 0.0]) if abs(z[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    x = np.cross(tmp, z); x /= np.linalg.norm(x)
    q = rng.uniform(0, 2 * math.pi)
    y = np.cross(z, x)
    x = math.cos(q) * x + math.sin(q) * y
    x -= np.dot(x, z) * z
    x /= np.linalg.norm(x)
    w = z * (2 * math.pi / (period_h * 3600.0)) if period_h else np.zeros(3)
    return x, z, w


def _radius_km(mass, rho):
    return (3 * mass / (4 * math.pi * rho)) ** (1 / 3) / 1000.0


def _loguni(rng, lo, hi):
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


def _barycentre(bodies: Sequence[Body]):
    gm = np.array([b.GM for b in bodies])
    r = np.array([b.r for b in bodies])
    v = np.array([b.v for b in bodies])
    rc, vc = (gm[:, None] * r).sum(0) / gm.sum(), (gm[:, None] * v).sum(0) / gm.sum()
    for b in bodies:
        b.r = b.r - rc
        b.v = b.v - vc


def _deformation_j2(b: Body, parent: Optional[Body] = None, r_parent: float = 0.0) -> float:
    """J2 of the rotational (k2r) and parent-tidal (k2) deformation that the
    loader removes from the input J2 (mass_loader.cpp:540-551)."""
    R = abs(b.R_km) * 1000.0
    w2 = float(np.dot(b.w, b.w))
    j2 = b.k2r * w2 * R ** 3 / (3 * b.GM)
    if parent is not None and r_parent > 0:
        j2 += b.k2 * parent.GM * R ** 3 / (2 * b.GM * r_parent ** 3)
    return j2


def _star(rng, tides):
    x, z, w = _spin_frame(rng, np.array([0.0, 0.0, 1.0]), 609.12, 0.13)
    s = Body("Star", "S0", G * M_SUN, 696000.0, inertia=0.0688, J2=2e-7, x_axis=x, z_axis=z, w=w)
    if tides:
        s.k2, s.k2r, s.td = 0.03, 0.027, 0.1
        s.J2 = 2e-7 + 1.5 * _deformation_j2(s)
    return s


def _hill_spaced(a_sorted, mass_sorted, ecc_sorted, m_central, k_hill=10.0):
    """True when neighbouring orbits are at least k_hill mutual Hill radii
    apart beyond their eccentric excursions (no close encounters)."""
    for q in range(len(a_sorted) - 1):
        a1, a2 = a_sorted[q], a_sorted[q + 1]
        rh = ((mass_sorted[q] + mass_sorted[q + 1]) / (3 * m_central)) ** (1 / 3) * (a1 + a2) / 2
        if a2 * (1 - ecc_sorted[q + 1]) - a1 * (1 + ecc_sorted[q]) < k_hill * rh:
            return False
    return True


def _planets(rng, star, n=8, tides=False, spaced=False):
    out = []
    if spaced:
        # orbits drawn until every neighbouring pair is Hill-stable
        while True:
            a_all = np.array([_loguni(rng, 0.39, 30.0) for _ in range(n)]) * AU
            e_all = rng.uniform(0, 0.1, n)
            m_all = np.array([_loguni(rng, 3e23, 2e27) for _ in range(n)])
            o = np.argsort(a_all)
            if _hill_spaced(a_all[o], m_all[o], e_all[o], M_SUN):
                break
    for k in range(n):
        a = _loguni(rng, 0.39, 30.0) * AU if not spaced else a_all[k]
        e = rng.uniform(0, 0.1) if not spaced else e_all[k]
        inc = rng.uniform(0, 3.0) * math.pi / 180
        mass = _loguni(rng, 3e23, 2e27) if not spaced else m_all[k]
        rho = rng.uniform(700, 5500) if mass > 1e26 else rng.uniform(3000, 5500)
        Om, om, M = rng.uniform(0, 2 * math.pi, 3)
        r, v, nrm = kepler_state(star.GM + G * mass, a, e, inc, Om, om, M)
        x, z, w = _spin_frame(rng, nrm, rng.uniform(9.5, 40.0), 0.5)
        b = Body(f"Planet {k + 1}", f"P{k + 1}", G * mass, _radius_km(mass, rho),
                 inertia=rng.uniform(0.22, 0.38), J2=rng.uniform(1e-3, 1.5e-2),
                 r=r, v=v, x_axis=x, z_axis=z, w=w)
        if tides:
            b.k2 = rng.uniform(0.1, 0.5)
            b.k2r = b.k2
            b.td = rng.uniform(0.0, 600.0)
            b.J2 += 1.5 * _deformation_j2(b, star, a)
        b._a, b._mass = a, mass
        out.append(b)
    return out


def _spaced_moon_orbits(rng, planet, ratios, lo, cap, k_hill=10.0):
    """Semi-major axes from lo outward, each pair of neighbours at least
    k_hill mutual Hill radii apart (beyond eccentricities <= 0.02); masses are
    scaled down by ten until the moons fit inside cap."""
    ratios = np.array(ratios, float)
    while True:
        a = [lo * rng.uniform(1.0, 1.3)]
        for q in range(1, len(ratios)):
            h = ((ratios[q - 1] + ratios[q]) / 3) ** (1 / 3)
            a.append(a[-1] * (1 + k_hill * h + 0.05) * rng.uniform(1.0, 1.15))
        if a[-1] <= cap or ratios.max() < 1e-12:
            return np.array(a), ratios
        ratios = ratios / 10


def _moons(rng, planet, count, prefix, major_ratio=None, tides=False, rmin=5.0, rmax=60.0,
           ratio_lo=1e-8, ratio_hi=1e-2, hill_cap=True, star_gm=G * M_SUN, spaced=False):
    out = []
    Rp = planet.R_km * 1000.0
    a_p = getattr(planet, "_a", None)
    if spaced and count:
        cap = 0.3 * a_p * (planet.GM / (3 * star_gm)) ** (1 / 3) if a_p else rmax * Rp
        lo = min(rmin * Rp, 0.1 * cap)
        ratios = [_loguni(rng, ratio_lo, ratio_hi) for _ in range(count)]
        rng.shuffle(ratios)
        a_sp, ratios = _spaced_moon_orbits(rng, planet, ratios, lo, cap)
    for q in range(count):
        ratio = _loguni(rng, ratio_lo, ratio_hi) if not spaced else ratios[q]
        mass = ratio * planet.GM / G
        hi = rmax * Rp
        if hill_cap and a_p:
            hi = min(hi, 0.3 * a_p * (planet.GM / (3 * star_gm)) ** (1 / 3))
        lo = min(rmin * Rp, 0.5 * hi)
        a = rng.uniform(lo, hi) if not spaced else a_sp[q]
        e = rng.uniform(0, 0.05) if not spaced else rng.uniform(0, 0.02)
        inc = rng.uniform(0, 0.1)
        Om, om, M = rng.uniform(0, 2 * math.pi, 3)
        r, v, nrm = kepler_state(planet.GM + G * mass, a, e, inc, Om, om, M)
        m = Body(f"{planet.name} moon {q + 1}", f"{prefix}{q + 1:02d}", G * mass,
                 _radius_km(mass, rng.uniform(1500, 3500)), r=planet.r + r, v=planet.v + v)
        if major_ratio is not None and ratio > major_ratio:
            period_h = 2 * math.pi * math.sqrt(a ** 3 / (planet.GM + G * mass)) / 3600.0
            m.x_axis, m.z_axis, m.w = _spin_frame(rng, nrm, period_h, 0.02)
            m.J2 = rng.uniform(1e-4, 1e-3)
            m.inertia = rng.uniform(0.3, 0.4)
            if tides:
                m.k2 = rng.uniform(0.02, 0.3)
                m.k2r = m.k2
                m.td = rng.uniform(0.0, 600.0)
                m.J2 += 1.5 * _deformation_j2(m, planet, a * (1 - e))
        out.append(m)
    return out


def make_small(seed: int = 20250) -> SystemSpec:
    """25-body star + planets + moons, Newtonian+1PN+J2."""
    rng = np.random.default_rng(seed)
    star = _star(rng, tides=False)
    planets = _planets(rng, star, 8, tides=False)
    bodies = [star]
    for k, p in enumerate(planets):
        bodies.append(p)
        bodies += _moons(rng, p, 2, f"M{k + 1}_", hill_cap=False)
    _barycentre(bodies)
    return SystemSpec(bodies, delta_t=300.0)


def _zonal_terms(rng, j2, nmax=6):
    return [(n, 0, j2 * rng.uniform(-0.05, 0.05) * 10.0 ** (-(n - 2)), 0.0) for n in range(3, nmax + 1)]


def _asteroid(rng, star: Body, q: int) -> Body:
    """Main-belt asteroid number q + 1: a uniform in 2.1-3.3 AU, e and i
    Rayleigh (sigma 0.1), mass log-uniform in 1e15-1e21 kg, spinning."""
    mass = _loguni(rng, 1e15, 1e21)
    a = rng.uniform(2.1, 3.3) * AU
    e = min(rng.rayleigh(0.1), 0.6)
    inc = min(rng.rayleigh(0.1), 1.2)
    Om, om, M = rng.uniform(0, 2 * math.pi, 3)
    r, v, nrm = kepler_state(star.GM + G * mass, a, e, inc, Om, om, M)
    x, z, w = _spin_frame(rng, rng.normal(size=3), rng.uniform(2.2, 30.0), 0.0)
    return Body(f"Asteroid {q + 1}", f"A{q + 1:04d}", G * mass, _radius_km(mass, 2200.0),
                r=r, v=v, x_axis=x, z_axis=z, w=w)


def make_full(seed: int = 20251) -> SystemSpec:
    """504-body system with the full force model."""
    rng = np.random.default_rng(seed)
    star = _star(rng, tides=True)
    star.gp, star.gp_terms = 1.0, _zonal_terms(rng, 2e-7)
    star.extra = {"MassRate": -0.3784717910405785, "Luminosity": 3.828e26}
    planets = _planets(rng, star, 8, tides=True, spaced=True)
    for p in planets:
        p.gp, p.gp_terms = 1.0, _zonal_terms(rng, p.J2)
    big = sorted(planets, key=lambda p: -p.GM)
    ringed = big[1]
    ringed.ring = 1.0
    Rk = ringed.R_km
    gm_ring = ringed.GM * 6.5e-8
    edges = [1.24, 1.53, 1.95, 2.03, 2.27]
    frac = [0.05, 0.70, 0.02, 0.23]
    ringed.rings = [(nm, edges[q] * Rk, edges[q + 1] * Rk, 2.0, gm_ring * frac[q])
                    for q, nm in enumerate(["C", "B", "Cassini", "A"])]
    planets[2].extra = {"Receptance": 1.457, "J2Rate": -1.55772e-18}
    planets[3].extra = {"Receptance": 1.2}
    w_moon = np.sqrt(np.array([p.GM for p in planets]) ** (1 / 3))
    alloc = rng.multinomial(77, w_moon / w_moon.sum())
    bodies = [star]
    for k, p in enumerate(planets):
        bodies.append(p)
        bodies += _moons(rng, p, int(alloc[k]), f"M{k + 1}_", major_ratio=1e-6, tides=True,
                         rmin=4.0, rmax=80.0, ratio_hi=1e-3, spaced=True)
    # dwarf planets: one in the main belt (also one of the 343 asteroids), four TNOs
    dwarfs = []
    dwarf_a = [rng.uniform(2.6, 2.9)] + list(rng.uniform(32, 48, 4))
    for k, a_au in enumerate(dwarf_a):
        mass = 9.4e20 if k == 0 else _loguni(rng, 1e21, 1.6e22)
        Om, om, M = rng.uniform(0, 2 * math.pi, 3)
        r, v, nrm = kepler_state(star.GM + G * mass, a_au * AU, rng.uniform(0, 0.25),
                                 rng.uniform(0, 0.3), Om, om, M)
        x, z, w = _spin_frame(rng, nrm, rng.uniform(6, 160), 0.6)
        d = Body(f"Dwarf {k + 1}", f"D{k + 1}", G * mass, _radius_km(mass, 2000.0),
                 inertia=0.33, J2=rng.uniform(1e-4, 5e-3), r=r, v=v, x_axis=x, z_axis=z, w=w)
        d._a = a_au * AU
        dwarfs.append(d)
    dm_alloc = [0, 1, 2, 5, 0]
    for k, d in enumerate(dwarfs):
        bodies.append(d)
        bodies += _moons(rng, d, dm_alloc[k], f"D{k + 1}_", rmin=5.0, rmax=60.0,
                         ratio_lo=1e-5, ratio_hi=1e-1, spaced=True)
    for q in range(342):
        ast = _asteroid(rng, star, q)
        if q < 3:
            ast.extra = {"Receptance": 1.0}
        bodies.append(ast)
    for q in range(27):
        mass = _loguni(rng, 1e17, 1e21)
        a = rng.uniform(30.0, 50.0) * AU
        e, inc = min(rng.rayleigh(0.1), 0.5), min(rng.rayleigh(0.1), 1.0)
        Om, om, M = rng.uniform(0, 2 * math.pi, 3)
        r, v, nrm = kepler_state(star.GM + G * mass, a, e, inc, Om, om, M)
        x, z, w = _spin_frame(rng, rng.normal(size=3), rng.uniform(4.0, 40.0), 0.0)
        bodies.append(Body(f"TNO {q + 1}", f"T{q + 1:03d}", G * mass, _radius_km(mass, 1500.0),
                           r=r, v=v, x_axis=x, z_axis=z, w=w))
    ring_mass = 0.02 * 5.97e24 / 36
    for q in range(36):
        lon = 2 * math.pi * q / 36
        r, v, _ = kepler_state(star.GM, 43.0 * AU, 0.0, 0.0, 0.0, 0.0, lon)
        bodies.append(Body(f"Kuiper ring {q + 1}", f"K{q + 1:02d}", G * ring_mass, -1.0e8,
                           r=r, v=v))
    assert len(bodies) == 504, len(bodies)
    _barycentre(bodies)
    return SystemSpec(bodies, delta_t=300.0)


def add_asteroids(spec: SystemSpec, count: int, seed: int = 20254) -> SystemSpec:
    """The system with `count` more main-belt asteroids (numbered after the
    existing ones; positions relative to the first body, which is the star),
    re-centred on the barycentre: systems larger than the 504-body one."""
    rng = np.random.default_rng(seed)
    bodies = [replace(b) for b in spec.bodies]
    star = bodies[0]
    n_ast = sum(b.sid.startswith("A") for b in bodies)
    for q in range(n_ast, n_ast + count):
        b = _asteroid(rng, star, q)
        b.r = b.r + star.r
        b.v = b.v + star.v
        bodies.append(b)
    _barycentre(bodies)
    return SystemSpec(bodies, spec.delta_t, spec.tdb)


def truncate(spec: SystemSpec, n: int) -> SystemSpec:
    """Keep the n most massive bodies, in their original order."""
    gm = np.array([b.GM for b in spec.bodies])
    keep = np.sort(np.argsort(-gm, kind="stable")[:n])
    return SystemSpec([replace(spec.bodies[k]) for k in keep], spec.delta_t, spec.tdb)


def perturb(spec: SystemSpec, rng, rel: float = 1e-9) -> SystemSpec:
    out = []
    for b in spec.bodies:
        c = replace(b)
        c.r = b.r * (1 + rel * rng.normal(size=3))
        c.v = b.v * (1 + rel * rng.normal(size=3))
        out.append(c)
    return SystemSpec(out, spec.delta_t, spec.tdb)


def make_ensemble(n_clones: int = 256, seed: int = 20252, base: Optional[SystemSpec] = None,
                  n_lo: int = 64, n_hi: int = 504) -> List[SystemSpec]:
    base = base or make_full()
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_lo, n_hi + 1, size=n_clones)
    return [perturb(truncate(base, int(nc)), rng) for nc in sizes]


def make_field9(base: Optional[SystemSpec] = None) -> SystemSpec:
    """Star (J2 only) and 8 point-mass planets of the full system."""
    base = base or make_full()
    star = replace(base.bodies[0], k2=0.0, k2r=0.0, td=0.0, gp=0.0, gp_terms=None, extra={})
    planets = [replace(b, k2=0.0, k2r=0.0, td=0.0, J2=0.0, gp=0.0, gp_terms=None, ring=0.0,
                       rings=None, extra={}, w=np.zeros(3))
               for b in base.bodies if b.sid.startswith("P")]
    bodies = [star] + planets
    _barycentre(bodies)
    return SystemSpec(bodies, delta_t=DAY)


PARTICLE_GM = 1e-20      # > eps^2 (sanity bound, mass_loader.cpp:33), negligible on every sum
PARTICLE_R_KM = 1.0


def make_particles(n: int, seed: int = 20253, field9: Optional[SystemSpec] = None):
    """Asteroid-like massless particles: (r[n,3], v[n,3]) float64."""
    field9 = field9 or make_field9()
    rng = np.random.default_rng(seed)
    mu = field9.bodies[0].GM
    a = rng.uniform(1.8, 3.5, n) * AU
    e = np.minimum(rng.rayleigh(0.12, n), 0.8)
    inc = np.minimum(rng.rayleigh(0.1, n), 1.2)
    Om, om, M = (rng.uniform(0, 2 * np.pi, n) for _ in range(3))
    E = M.copy()
    for _ in range(50):
        E -= (E - e * np.sin(E) - M) / (1 - e * np.cos(E))
    cE, sE = np.cos(E), np.sin(E)
    rad = a * (1 - e * cE)
    xo, yo = a * (cE - e), a * np.sqrt(1 - e * e) * sE
    fac = np.sqrt(mu * a) / rad
    vxo, vyo = -fac * sE, fac * np.sqrt(1 - e * e) * cE
    cO, sO, ci, si, co, so = np.cos(Om), np.sin(Om), np.cos(inc), np.sin(inc), np.cos(om), np.sin(om)
    P = np.stack([cO * co - sO * so * ci, sO * co + cO * so * ci, so * si], -1)
    Q = np.stack([-cO * so - sO * co * ci, -sO * so + cO * co * ci, co * si], -1)
    star = field9.bodies[0]
    r = xo[:, None] * P + yo[:, None] * Q + star.r
    v = vxo[:, None] * P + vyo[:, None] * Q + star.v
    return np.ascontiguousarray(r), np.ascontiguousarray(v)


def particle_body(r, v, sid="X0") -> Body:
    """A massless test particle as an Initial-file row (used to cross-check
    the particle kernel against a full system holding one such body)."""
    return Body("Particle", sid, PARTICLE_GM, PARTICLE_R_KM, r=np.asarray(r, float),
                v=np.asarray(v, float))


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("workload", choices=["small", "full", "field9"])
    ap.add_argument("out_dir")
    ap.add_argument("--seed", type=int, default=None)
    a = ap.parse_args(argv)
    fn = {"small": make_small, "full": make_full, "field9": make_field9}[a.workload]
    spec = fn() if a.seed is None or a.workload == "field9" else fn(a.seed)
    print(write_system(spec, a.out_dir, a.workload.capitalize()))


if __name__ == "__main__":
    main()
