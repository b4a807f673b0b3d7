"""ctypes mirror of include/mcsim.h (the C ABI of libmcsim.so).

Only plain data layouts live here; loading the library is done by ``mcsim.lib``.
"""
import ctypes as C

MC_BINS = 180
MC_MAX_MODES = 4
MC_MAX_FLUTES = 8
MC_MAX_SLICES = 64
MC_CHATTER_IDX = 0.45

MC_OK, MC_ERR_ARG, MC_ERR_LIMIT, MC_ERR_NODEV, MC_ERR_HIP, MC_ERR_NOMEM = range(6)
MC_TOOL_FLAT, MC_TOOL_BALL, MC_TOOL_BULL, MC_TOOL_PROFILE = 0, 1, 2, 3


class Tool(C.Structure):
    _fields_ = [("N", C.c_int32), ("type", C.c_int32), ("R", C.c_double),
                ("helixDeg", C.c_double), ("fluteLen", C.c_double),
                ("runout", C.c_double), ("runoutPhase", C.c_double)]


class Mode(C.Structure):
    _fields_ = [("k", C.c_double), ("fn", C.c_double), ("zeta", C.c_double)]


class Material(C.Structure):
    _fields_ = [("Ktc", C.c_double), ("Krc", C.c_double), ("Kac", C.c_double),
                ("Kte", C.c_double), ("Kre", C.c_double), ("Kae", C.c_double),
                ("mc", C.c_double), ("Ci", C.c_double)]


class Cell(C.Structure):
    _fields_ = [("tool", Tool), ("mat", Material), ("nmx", C.c_int32), ("nmy", C.c_int32),
                ("mx", Mode * MC_MAX_MODES), ("my", Mode * MC_MAX_MODES),
                ("S", C.c_double), ("phiSt", C.c_double), ("phiEx", C.c_double),
                ("fz", C.c_double), ("revs", C.c_int32), ("nonlinear", C.c_int32)]


class Trial(C.Structure):
    _fields_ = [("index", C.c_double), ("amp", C.c_double), ("once", C.c_double),
                ("chatter", C.c_int32), ("nz", C.c_int32), ("steps", C.c_int64)]


class BisectOpts(C.Structure):
    _fields_ = [("aMin", C.c_double), ("aMax", C.c_double), ("iters", C.c_int32),
                ("reserved", C.c_int32)]


class Limit(C.Structure):
    _fields_ = [("alim", C.c_double), ("lo", C.c_double), ("hi", C.c_double),
                ("open", C.c_int32), ("trials", C.c_int32), ("steps", C.c_int64),
                ("visits", C.c_int64), ("engaged", C.c_int64)]


class Stamp(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("tipZ", C.c_double),
                ("R", C.c_double), ("type", C.c_int32), ("line", C.c_int32),
                ("errVal", C.c_double)]


class StampOpt(C.Structure):
    _fields_ = [("ux", C.c_double), ("uy", C.c_double), ("f", C.c_double),
                ("lut", C.c_int32), ("pad", C.c_int32)]


class EngageSpec(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("ux", "uy", "fz", "tipZ", "ap", "vx0", "vy0", "fnMax")] + \
               [("Hb", C.c_float * 180)]


class EngageQuery(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("cx", "cy", "ux", "uy", "R", "tipZ", "fluteLen")]


class EngageResult(C.Structure):
    _fields_ = [("hmax", C.c_double), ("ap", C.c_double), ("phiSt", C.c_double),
                ("phiEx", C.c_double), ("imm", C.c_int32), ("pad", C.c_int32)]


MC_END_BINS = 36
MC_END_RINGS = 6
MC_ENGINE_HIST = 4096
MC_ENGINE_RING = 64


class EngineInfo(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dt", "omega", "Ttooth", "dz", "S")] + \
               [(n, C.c_int32) for n in ("stepsPerTooth", "stepsPerRev", "decim", "piStep", "maxSlices",
                                         "nmx", "nmy", "pad")]


class Wall(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("k", "m", "zeta", "nx", "ny")]


class EngineState(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "vx", "vy", "w", "wv", "t", "runMeanW", "wk", "wnx",
                                          "wny")] + \
               [("stepCount", C.c_int64)] + [(n, C.c_int32) for n in ("pr", "decim", "nz", "pad")] + \
               [(n, C.c_int64) for n in ("hist_i", "hist_count", "ts_i", "ts_count", "ws_i", "ws_count",
                                         "wsp_i", "wsp_count")]


STATS_FIELDS = ("n", "sumF", "maxF", "sumFx", "sumFy", "sumFz", "maxFz", "sumT", "maxT", "sumX", "sumY",
                "minX", "maxX", "minY", "maxY", "sumZc", "wZc", "sumW", "maxW", "sumHb", "sumB", "sumRb",
                "engTeeth")


class EngineStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in STATS_FIELDS]


class FrameOut(C.Structure):
    _fields_ = [("stats", EngineStats), ("index", C.c_double), ("amp", C.c_double), ("once", C.c_double),
                ("end", C.c_int32), ("pad", C.c_int32)]


class ChipEvent(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("vol", "phi", "z", "ux", "uy", "vx", "vy")]


class Substep(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "ux", "uy", "dtm", "fzEff", "fza", "dxy")] + \
               [("rapid", C.c_int32), ("line", C.c_int32), ("nSubEst", C.c_double), ("guard", C.c_int32),
                ("frame", C.c_int32)]


MC_FRAME_FIRST, MC_FRAME_LAST = 1, 2


class Move(C.Structure):
    _fields_ = [("from_", C.c_double * 3), ("to", C.c_double * 3), ("F", C.c_double), ("S", C.c_double),
                ("rapid", C.c_int32), ("line", C.c_int32), ("arc", C.c_int32), ("cw", C.c_int32),
                ("axes", C.c_int32 * 3), ("has_sweep", C.c_int32), ("ca", C.c_double), ("cb", C.c_double),
                ("sweep", C.c_double)]


class PathPoint(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "F", "S")] + [("rapid", C.c_int32), ("line", C.c_int32),
                                                                         ("move", C.c_int32), ("pad", C.c_int32)]


class ScheduleOpts(C.Structure):
    _fields_ = [("hx", C.c_double), ("dt_real", C.c_double), ("speed", C.c_double), ("feed_override", C.c_double),
                ("flutes", C.c_int32), ("max_frames", C.c_int32)]


class SubstepOut(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("dx", "dy", "ap", "errVal", "vol", "maxF", "phiSt", "phiEx")] + \
               [("nsteps", C.c_int32), ("cut", C.c_int32), ("arc", C.c_int32), ("pad", C.c_int32)]


WALL_FIELDS = ("k", "m", "fn", "zeta", "nx", "ny", "t", "tRoot", "hWall", "a", "bEff", "base", "cx", "cy", "I")


class WallOut(C.Structure):
    """mc_wall_out: the wall wallAt found at a sub-step and the wall-error mark after it"""
    _fields_ = [(n, C.c_double) for n in WALL_FIELDS + ("markVal", "wnx", "wny")] + \
               [("found", C.c_int32), ("marked", C.c_int32)]


MC_WALL_FN_CAP = 5000


class SimProgram(C.Structure):
    _fields_ = [("actual", C.c_void_p), ("ideal", C.c_void_p), ("engine", C.c_void_p),
                ("sub", C.POINTER(Substep)), ("out", C.POINTER(SubstepOut)),
                ("R", C.c_double), ("fluteLen", C.c_double), ("rw", C.c_double), ("share", C.c_double),
                ("nsub", C.c_int32), ("type", C.c_int32), ("lut", C.POINTER(C.c_float)), ("lutn", C.c_int32),
                ("nframes", C.c_int32), ("cornerR", C.c_double), ("dynBudget", C.c_double), ("fz", C.c_double),
                ("frames", C.c_void_p), ("flex", C.c_int32), ("pad", C.c_int32), ("wallE", C.c_double),
                ("wallRho", C.c_double), ("kTool", C.c_double), ("walls", C.POINTER(WallOut)),
                ("edge_r", C.POINTER(C.c_double)), ("edge_dz", C.c_double), ("edge_n", C.c_int32),
                ("pad3", C.c_int32), ("budgetLeft", C.POINTER(C.c_double))]


MC_STOCK_B = 8
MC_LUT_DR = 0.02
MC_TIP_FLAT, MC_TIP_BALL, MC_TIP_BULL, MC_TIP_EDGE = 0, 1, 2, 3


def make_cell(tool, modes_x, modes_y, mat, S, arc, fz, revs, nonlinear=True):
    """Build a Cell from plain values.

    tool: dict(N, R, helixDeg, type ('flat'|'ball' or int), fluteLen, runout=0, runoutPhase=0)
    modes_x / modes_y: lists of (k [N/mm], fn [Hz], zeta); modes_y None -> modes_x
    mat: dict with Ktc Krc Kac Kte Kre Kae mc Ci (missing -> 0)
    """
    c = Cell()
    t = tool.get("type", 0)
    c.tool = Tool(int(tool["N"]), MC_TOOL_BALL if t in ("ball", 1) else MC_TOOL_FLAT,
                  float(tool["R"]), float(tool["helixDeg"]), float(tool["fluteLen"]),
                  float(tool.get("runout", 0.0)), float(tool.get("runoutPhase", 0.0)))
    c.mat = Material(*[float(mat.get(k, 0.0)) for k in
                       ("Ktc", "Krc", "Kac", "Kte", "Kre", "Kae", "mc", "Ci")])
    modes_y = modes_x if not modes_y else modes_y
    if not (1 <= len(modes_x) <= MC_MAX_MODES and 1 <= len(modes_y) <= MC_MAX_MODES):
        raise ValueError("1..%d modes per axis" % MC_MAX_MODES)
    c.nmx, c.nmy = len(modes_x), len(modes_y)
    for i, m in enumerate(modes_x):
        c.mx[i] = Mode(*map(float, m))
    for i, m in enumerate(modes_y):
        c.my[i] = Mode(*map(float, m))
    c.S = float(S)
    c.phiSt, c.phiEx = float(arc[0]), float(arc[1])
    c.fz = float(fz)
    c.revs = int(revs)
    c.nonlinear = 1 if nonlinear else 0
    return c


def cell_array(cells):
    arr = (Cell * len(cells))()
    for i, c in enumerate(cells):
        arr[i] = c
    return arr


MC_SURR_IN = 24
MC_SURR_HIDDEN = 256


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float),
                ("lr_final", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("fit_scaling", C.c_int32), ("seed", C.c_uint64)]


class LobeArgs(C.Structure):
    _fields_ = [("N", C.c_int32), ("lobes", C.c_int32), ("Ktc", C.c_double), ("Krc", C.c_double),
                ("phiSt", C.c_double), ("phiEx", C.c_double), ("rpmMin", C.c_double),
                ("rpmMax", C.c_double), ("nmx", C.c_int32), ("nmy", C.c_int32),
                ("mx", Mode * MC_MAX_MODES), ("my", Mode * MC_MAX_MODES)]


class SearchTuning(C.Structure):
    _fields_ = [("waves_per_cu", C.c_int32), ("wide_max", C.c_int32), ("wide_frac", C.c_double),
                ("use_graph", C.c_int32), ("reserved", C.c_int32 * 3)]
