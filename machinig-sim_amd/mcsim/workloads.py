"""Workload definitions for the two hot paths, with their exact sizes.

* ``sld_40rpm()``     40-speed stability lobe diagram, one tool, one mode per axis
  (3-flute 12 mm end mill, 30 deg helix, fn 1000 Hz, zeta 0.03, k 2e4 N/mm,
  half-immersion down-milling, 40 revolutions per trial).  FP64, CPU-runnable.
* ``sld_grid()``      256 speeds x 64 seeded tool/mode variants (16,384 cells),
  2-4 flutes, 1-4 modes per axis, immersion 10-100 %, process damping on,
  20-80 revolutions per trial.
* ``zigzag_path()``   512 x 512 heightmap at 0.1 mm, 5,000 seeded zigzag G1 segments
  alternating a flat 10 mm and a ball 6 mm tool, and the stamp schedule of
  src/sim.js:355-377 along it.

Random draws use the Park-Miller generator of the reference's tests
(seed <- seed * 16807 mod 2147483647, test/identify.test.mjs:20).
Units follow the reference: modal k in N/mm (src/cutting.js:55), lengths in mm.
"""
import math

from ._abi import make_cell, MC_TOOL_BALL, MC_TOOL_FLAT

JS_PI = 3.141592653589793


class ParkMiller:
    def __init__(self, seed=1):
        self.seed = int(seed)

    def u(self):
        """uniform in (0, 1)"""
        self.seed = (self.seed * 16807) % 2147483647
        return self.seed / 2147483647

    def uniform(self, a, b):
        return a + (b - a) * self.u()

    def log_uniform(self, a, b):
        return math.exp(math.log(a) + (math.log(b) - math.log(a)) * self.u())

    def integer(self, a, b):
        """uniform integer in [a, b]"""
        return a + min(b - a, int(self.u() * (b - a + 1)))


def arc_of(ae, D):
    """engagement arc of down milling at radial depth ae (src/conditions.js:11-15)"""
    if not (ae < D):
        return (0.0, JS_PI)
    return (JS_PI - math.acos(1 - 2 * ae / D), JS_PI)


# cutting coefficients: the 40-speed diagram's own set, and the A5052 row of
# src/materials.js:25 (with its size effect mc 0.10 and process damping Ci 2e4)
MAT_SLD40 = dict(Ktc=700.0, Krc=210.0, Kac=0.0, Kte=20.0, Kre=15.0, Kae=0.0, mc=0.0, Ci=0.0)
MAT_A5052 = dict(Ktc=700.0, Krc=210.0, Kac=120.0, Kte=20.0, Kre=25.0, Kae=5.0, mc=0.10, Ci=2e4)

BISECT_DEFAULT = dict(aMin=0.05, aMax=10.0, iters=10)


def sld_40rpm():
    """40 cells: S_i = 3000 + 12000 i / 39, i = 0..39"""
    tool = dict(N=3, R=6.0, helixDeg=30.0, type="flat", fluteLen=24.0)
    modes = [(2e4, 1000.0, 0.03)]
    arc = (JS_PI / 2, JS_PI)
    cells = []
    for i in range(40):
        S = 3000 + 12000 * i / 39
        cells.append(make_cell(tool, modes, None, MAT_SLD40, S, arc, 0.05, 40, True))
    return cells, dict(BISECT_DEFAULT)


def grid_tools(n_tools=64, seed=1):
    """64 seeded tool / mode-set / immersion variants"""
    rng = ParkMiller(seed)
    tools = []
    for _ in range(n_tools):
        N = rng.integer(2, 4)
        helix = rng.uniform(15, 45)
        D = rng.uniform(6, 20)
        runout = rng.uniform(0, 0.010)          # mm; chatterAt forces 0 (src/stabtd.js:16)
        nmx = rng.integer(1, 4)
        nmy = rng.integer(1, 4)
        mx = [(rng.log_uniform(5e3, 5e4), rng.log_uniform(500, 4000), rng.uniform(0.01, 0.05))
              for _ in range(nmx)]
        my = [(rng.log_uniform(5e3, 5e4), rng.log_uniform(500, 4000), rng.uniform(0.01, 0.05))
              for _ in range(nmy)]
        imm = rng.uniform(0.1, 1.0)
        revs = rng.integer(20, 80)
        tools.append(dict(tool=dict(N=N, R=D / 2, helixDeg=helix, type="flat", fluteLen=24.0,
                                    runout=runout),
                          mx=mx, my=my, imm=imm, D=D, revs=revs))
    return tools


def grid_speeds(n=256):
    return [2000 + 28000 * i / (n - 1) for i in range(n)]


def sld_grid(n_speeds=256, n_tools=64, seed=1):
    """n_tools x n_speeds cells (tool-major).  Defaults: 16,384 cells."""
    tools = grid_tools(n_tools, seed)
    speeds = grid_speeds(n_speeds)
    cells = []
    for t in tools:
        arc = arc_of(t["imm"] * t["D"], t["D"])
        for S in speeds:
            cells.append(make_cell(t["tool"], t["mx"], t["my"], MAT_A5052, S, arc, 0.05,
                                   t["revs"], True))
    return cells, dict(BISECT_DEFAULT)


# ---------------------------------------------------------------- heightmap path


def zigzag_segments(n_segments=5000, width=51.1, seed=1):
    """Seeded zigzag G1 program.  Layers alternate flat R5 / ball R3; per layer a depth
    d ~ U(0.5, 3) (tip at z = -d) and a stepover ~ U(2, 5); raster lines along x across
    the stock, stepover moves along y, a diagonal link to the next layer's start.
    Returns a list of (x0, y0, z0, x1, y1, z1, R, type, line)."""
    rng = ParkMiller(seed)
    segs = []
    layer = 0
    pos = (0.0, 0.0, 0.0)
    while len(segs) < n_segments:
        flat = layer % 2 == 0
        R, typ = (5.0, MC_TOOL_FLAT) if flat else (3.0, MC_TOOL_BALL)
        d = rng.uniform(0.5, 3.0)
        step = rng.uniform(2.0, 5.0)
        z = -d
        y = 0.0
        x_left = 0.0
        start = (x_left, y, z)
        if len(segs) < n_segments:
            segs.append(pos + start + (R, typ, len(segs) + 1))
        pos = start
        direction = 1
        while len(segs) < n_segments:
            x_end = width if direction > 0 else 0.0
            nxt = (x_end, pos[1], z)
            segs.append(pos + nxt + (R, typ, len(segs) + 1))
            pos = nxt
            if pos[1] + step > width or len(segs) >= n_segments:
                break
            nxt = (pos[0], pos[1] + step, z)
            segs.append(pos + nxt + (R, typ, len(segs) + 1))
            pos = nxt
            direction = -direction
        layer += 1
    return segs[:n_segments]


# the zigzag program run through the simulator's sub-step loop: a 10 mm flat and a 6 mm ball
# end mill (three flutes, 30 deg helix), one rigid mode per axis as the reference's target test
# (k 5e6 N/m, fn 6000 Hz, zeta 0.1, test/target.test.mjs:66), A5052, feed 1000 mm/min at
# 10000 min^-1 (fz = F / (S N))
ZIGZAG_TOOLS = {MC_TOOL_FLAT: dict(N=3, R=5.0, helixDeg=30.0, type="flat", fluteLen=20.0),
                MC_TOOL_BALL: dict(N=3, R=3.0, helixDeg=30.0, type="ball", fluteLen=20.0)}
ZIGZAG_MODES = [(5.0e3, 6000.0, 0.1)]
ZIGZAG_FEED, ZIGZAG_SPEED = 1000.0, 10000.0


def zigzag_layers(n_segments=5000, width=51.1, seed=1):
    """The zigzag program of zigzag_segments as the simulator runs it, one entry per layer
    (each layer has its own tool): dict(type, tool, x, y, z, rapid, line, F, S), the path's
    points (point 0 the position the layer starts from, point k the end of the layer's k-th
    segment, which carries that segment's program line), every segment a feed move."""
    layers = []
    for sg in zigzag_segments(n_segments, width, seed):
        x0, y0, z0, x1, y1, z1, R, typ, line = sg
        if not layers or layers[-1]["type"] != typ:
            layers.append(dict(type=typ, tool=ZIGZAG_TOOLS[typ], x=[x0], y=[y0], z=[z0], rapid=[1], line=[line],
                               F=[None], S=[ZIGZAG_SPEED]))
        L = layers[-1]
        L["x"].append(x1); L["y"].append(y1); L["z"].append(z1); L["rapid"].append(0); L["line"].append(line)
        L["F"].append(ZIGZAG_FEED); L["S"].append(ZIGZAG_SPEED)
    return layers


def zigzag_coupled(layers, stocks, engine, run, cell_of, sub_of, max_substeps=None):
    """Run the layers through the sub-step loop one after another on one pair of stocks
    (actual, commanded) and one engine, configured for each layer's tool as the simulator's
    ensureEngine does.  run(actual, ideal, engine, sub, n, tool, share) runs n sub-steps and
    returns their SubstepOut records; cell_of(tool) gives the engine's configure argument and
    sub_of(layer) the layer's sub-steps.  Stops after max_substeps.  Returns the records and
    each record's program line."""
    act, ide = stocks
    outs, lines = [], []
    left = max_substeps
    for L in layers:
        sub, n = sub_of(L)
        if left is not None:
            n = min(n, left)
            left -= n
        if n <= 0:
            break
        info = engine.configure(cell_of(L["tool"]))
        engine.set_wear(MAT_A5052["Kte"], MAT_A5052["Kre"])
        outs.extend(run(act, ide, engine, sub, n, L["tool"], info.stepsPerTooth / 4))
        lines.extend(sub[k].line for k in range(n))
    return outs, lines


def segment_max_force(outs, lines):
    """per program line (segment): the largest |(Fx, Fy)| over its cutting sub-steps' engine
    runs (the run statistics' maxF, src/cutting.js:280-281); 0 for a segment that never cut"""
    seg = {}
    for o, ln in zip(outs, lines):
        seg[ln] = max(seg.get(ln, 0.0), o.maxF)
    return seg


# ---------------------------------------------------------------- coupled sub-step programs


def coupled_program(seed, width, n_lines=4, feed=1000.0, speed=10000.0):
    """One seeded raster program for the simulator's coupled sub-step loop: a rapid to a
    start point beside the stock, a plunge at a third of the feed (end edges cut), then
    n_lines raster passes along x across the stock at a depth d ~ U(0.5, 3) with a stepover
    ~ U(2, 5) along y, alternating direction.  Returns the discretised path as lists
    (x, y, z, rapid, line, F, S), point 0 being the start of the first move."""
    rng = ParkMiller(seed)
    d = rng.uniform(0.5, 3.0)
    step = rng.uniform(2.0, 5.0)
    y = rng.uniform(4.0, 10.0)
    pts = [(-6.0, y, 5.0, 1), (-6.0, y, 2.0, 1), (2.0, y, 2.0, 1), (2.0, y, -d, 0)]
    F = [None, None, None, feed / 3]
    direction = 1
    x = 2.0
    for k in range(n_lines):
        x = width - 2.0 if direction > 0 else 2.0
        pts.append((x, y, -d, 0)); F.append(feed)
        if k + 1 < n_lines:
            y += step
            pts.append((x, y, -d, 0)); F.append(feed)
        direction = -direction
    xs = [p[0] for p in pts]
    ys = [p[1] for p in pts]
    zs = [p[2] for p in pts]
    rapid = [p[3] for p in pts]
    line = list(range(1, len(pts) + 1))
    S = [speed] * len(pts)
    return xs, ys, zs, rapid, line, F, S


def coupled_full_program(seed, width, n_segments=100000, feed=1000.0, speed=10000.0, chord=0.09):
    """One seeded mixed roughing / finishing program of about n_segments linear moves for the
    coupled sub-step loop on a width x width stock: a plunge beside the stock's corner, a
    zigzag pocket at depth d ~ U(0.5, 3) with stepover ~ U(2, 5) along y over the whole
    stock, then contour passes: rounded-rectangle loops (corner radius 8 mm) inset by
    ~ U(3, 6) mm from the stock's edge, each loop 0.2 mm deeper than the last below the
    pocket floor and 0.3 mm further in, chorded into moves of about `chord` mm (as the path
    discretizer chords arcs), until the program holds n_segments moves.  Returns the path as
    lists (x, y, z, rapid, line, F, S), point 0 being the start of the first move."""
    import math
    rng = ParkMiller(seed)
    d = rng.uniform(0.5, 3.0)
    step = rng.uniform(2.0, 5.0)
    inset = rng.uniform(3.0, 6.0)
    y = 2.0
    pts = [(-6.0, y, 5.0, 1), (-6.0, y, 2.0, 1), (2.0, y, 2.0, 1), (2.0, y, -d, 0)]
    F = [None, None, None, feed / 3]
    direction = 1
    while True:
        x = width - 2.0 if direction > 0 else 2.0
        pts.append((x, y, -d, 0)); F.append(feed)
        if y + step > width - 2.0:
            break
        y += step
        pts.append((x, y, -d, 0)); F.append(feed)
        direction = -direction
    # contour loops: a rounded rectangle traced counter-clockwise from the middle of its
    # lower edge, chorded; each loop starts with a short ramp to its depth
    rc, z, o = 8.0, -d, inset
    while len(pts) < n_segments + 1:
        z -= 0.2
        lo, hi = o, width - o
        if hi - lo < 2 * rc + 1.0:
            break
        path = []
        cx0, cy0 = (lo + hi) / 2, lo
        path.append((cx0, cy0))
        corners = [(hi - rc, lo + rc, -math.pi / 2), (hi - rc, hi - rc, 0.0), (lo + rc, hi - rc, math.pi / 2),
                   (lo + rc, lo + rc, math.pi)]
        edges = [((cx0, lo), (hi - rc, lo)), ((hi, lo + rc), (hi, hi - rc)), ((hi - rc, hi), (lo + rc, hi)),
                 ((lo, hi - rc), (lo, lo + rc))]
        for (a, b), (kx, ky, a0) in zip(edges, corners):
            L = math.hypot(b[0] - a[0], b[1] - a[1])
            m = max(1, int(math.ceil(L / chord)))
            for i in range(1, m + 1):
                path.append((a[0] + (b[0] - a[0]) * i / m, a[1] + (b[1] - a[1]) * i / m))
            m = max(1, int(math.ceil(rc * math.pi / 2 / chord)))
            for i in range(1, m + 1):
                t = a0 + math.pi / 2 * i / m
                path.append((kx + rc * math.cos(t), ky + rc * math.sin(t)))
        L = math.hypot(cx0 - lo - rc, 0.0)
        m = max(1, int(math.ceil(L / chord)))
        for i in range(1, m + 1):
            path.append((lo + rc + (cx0 - lo - rc) * i / m, lo))
        # rapid up, over to the loop's start, down, then ramp along the first chords
        pts.append((pts[-1][0], pts[-1][1], 2.0, 1)); F.append(None)
        pts.append((path[0][0], path[0][1], 2.0, 1)); F.append(None)
        pts.append((path[0][0], path[0][1], z + 0.2, 1)); F.append(None)
        for i, (px, py) in enumerate(path[1:]):
            if len(pts) >= n_segments + 1:
                break
            pz = z + 0.2 * max(0.0, 1.0 - (i + 1) / 20.0)
            pts.append((px, py, pz, 0)); F.append(feed)
        o += 0.3
    xs = [p[0] for p in pts]
    ys = [p[1] for p in pts]
    zs = [p[2] for p in pts]
    rapid = [p[3] for p in pts]
    line = list(range(1, len(pts) + 1))
    S = [speed] * len(pts)
    return xs, ys, zs, rapid, line, F, S


# the coupled programs' tools, modes and material: a 12 mm flat end mill and an 8 mm ball
# end mill taking turns, two modes per axis, A5052 with the size effect and process damping;
# program p runs n_lines - 1 + p % 4 passes, so the programs differ in length
COUPLED_TOOLS = [dict(N=3, R=6.0, helixDeg=30.0, type="flat", fluteLen=30.0, runout=0.005, runoutPhase=0.0),
                 dict(N=2, R=4.0, helixDeg=30.0, type="ball", fluteLen=20.0, runout=0.005, runoutPhase=0.0)]
COUPLED_MODES = [(2.0e4, 1800.0, 0.03), (6.0e4, 3200.0, 0.04)]
COUPLED_SPEED = 10000.0   # min^-1, every line of the coupled programs


def coupled_plan(p, n_lines):
    """(tool, passes) of coupled program p"""
    return COUPLED_TOOLS[p % 2], n_lines - 1 + p % 4
