"""ctypes binding of the CPU oracle (liboracle_softbox.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker.  The product path (softbox_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_softbox.so")
LIB_OMP_PATH = os.path.join(HERE, "liboracle_softbox_omp.so")   # same restatement, OpenMP per-item loops
LIB_F64_PATH = os.path.join(HERE, "liboracle_softbox_f64.so")   # every float of the state widened to double

vp = C.c_void_p
i32 = C.c_int
u32 = C.c_uint32


def build() -> str:
    """Compile the oracle, its OpenMP build and its FP64 build with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "all"], check=True)
    return LIB_PATH


FORM_SCRATCH_I = ("filID", "filIDScratch", "bodyCell", "cellCount", "gridCellOffsets", "gridCellContents",
                  "candCountSeg", "candBaseSeg", "reqFilA", "reqFilB", "reqOrient", "gatePass", "minCand",
                  "acceptFlag", "freeCount", "freeOffsets", "freeList", "rankOffsets")
FORM_SCRATCH_F = ("center", "reqLoc1", "reqLoc2")


class SboNet(C.Structure):
    _fields_ = (
        [("nRod", i32)]
        + [(k, vp) for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam",
                             "forceSum", "torqueSum", "randForce", "randTorque", "brownTransScale", "brownRotScale",
                             "end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide")]
        + [("params", C.c_float * 2), ("chainParams", C.c_float * 7), ("chainOn", i32), ("nLinks", i32)]
        + [(k, vp) for k in ("linkFilA", "linkFilB", "linkState", "strainPlace", "linkOrientSame", "torqueMagPlace",
                             "activeLinkCount", "loc1", "loc2", "strainHist", "torqueMagHist", "xlinkData")]
        + [("xlParams", C.c_float * 3), ("offParams", C.c_float * 5), ("torsionParams", C.c_float * 2),
           ("unbindOn", i32)]
        + [(k, vp) for k in ("segCountA", "segOffsetsA", "segIdxA", "segCountB", "segOffsetsB", "segIdxB")]
        + [("confineOn", i32), ("boxParams", C.c_float * 8), ("gridOn", i32), ("periodic", i32), ("maxCand", i32),
           ("gridParams", C.c_float * 6), ("boxLen", C.c_float * 3), ("actinRadius", C.c_float),
           ("gridDims", C.c_int * 4)]
        + [(k, vp) for k in ("center", "boundingRadius", "bodyCell", "cellCount", "gridCellOffsets",
                             "gridCellContents", "candPartner", "candCount")]
        + [("seed", i32), ("rng", i32), ("stream", u32), ("form", vp), ("formInt", i32), ("turn", vp)]
    )


class SboForm(C.Structure):
    _fields_ = (
        [("nSeg", i32), ("C", i32), ("reqCap", i32)]
        + [(k, vp) for k in ("coord", "uVec", "end1", "end2", "segLength", "filState", "end2NbrSlot", "linkFilA",
                             "linkFilB", "linkState", "strainPlace", "linkOrientSame", "torqueMagPlace",
                             "activeLinkCount", "loc1", "loc2", "strainHist", "torqueMagHist")]
        + [("gridParams", C.c_float * 6), ("gridDims", C.c_int * 4), ("periodic", i32), ("boxLen", C.c_float * 3),
           ("formParams", C.c_float * 7), ("mode", i32), ("formCounts", C.c_int * 4)]
        + [(k, vp) for k in FORM_SCRATCH_I + FORM_SCRATCH_F]
        + [("lastFormed", i32)]
    )


TURN_I = ("filState", "monomerCount", "seedNode")
TURN_F = ("nucFrac", "cofFrac", "bTransDiff", "bRotDiff")
TURN_SCRATCH_I = ("grewFlag", "returnedMon", "deathFlag", "severDeathFlag", "severReturnedMon", "acceptFlag")
TURN_SCRATCH_F = ("reqCoord", "reqUVec", "reqYVec")
TURN_OFFS = ("freeCount", "freeOffsets", "freeList", "rankOffsets", "scanOffsets")


class SboTurnover(C.Structure):
    _fields_ = (
        [("C", i32)]
        + [(k, vp) for k in TURN_I + TURN_F + TURN_SCRATCH_I + TURN_SCRATCH_F + TURN_OFFS]
        + [("interval", i32), ("seed", i32), ("growthOn", i32), ("proxy", i32)]
        + [("growParams", C.c_float * 3), ("splitParams", C.c_float * 3), ("dragParams", C.c_float * 9),
           ("depolyParams", C.c_float * 3), ("depolyRateParams", C.c_float * 2), ("agingParams", C.c_float * 2),
           ("cofilinParams", C.c_float * 2), ("birthParams", C.c_float * 2)]
        + [("onRate", C.c_double), ("biochemDeltaT", C.c_double), ("uMPerMonomer", C.c_double), ("conc", C.c_double),
           ("taken", C.c_longlong), ("returned", C.c_longlong), ("last", C.c_int * 6), ("cadences", i32)]
    )


def _widen(fields):
    """The same struct with every float field and float array widened to double (the FP64 build)."""
    out = []
    for name, t in fields:
        if t is C.c_float:
            t = C.c_double
        elif getattr(t, "_type_", None) is C.c_float:
            t = C.c_double * t._length_
        out.append((name, t))
    return out


class SboNet64(C.Structure):
    _fields_ = _widen(SboNet._fields_)


class SboForm64(C.Structure):
    _fields_ = _widen(SboForm._fields_)


class SboTurnover64(C.Structure):
    _fields_ = _widen(SboTurnover._fields_)


def _fill(arr, values) -> None:
    """Copy parameters into a struct array; floats are rounded to float32 first, so that the FP32 and
    FP64 runs start from the same parameter values."""
    floating = arr._type_ in (C.c_float, C.c_double)
    for i, v in enumerate(values):
        arr[i] = float(np.float32(v)) if floating else v


def _widen_store(obj) -> None:
    """Every float32 array of a host store (and of the stores it holds) replaced by a float64 copy."""
    for k, v in list(vars(obj).items()):
        if isinstance(v, np.ndarray) and v.dtype == np.float32:
            setattr(obj, k, v.astype(np.float64))
        elif hasattr(v, "__dict__") and any(isinstance(x, np.ndarray) for x in vars(v).values()):
            _widen_store(v)


_lib = None
_lib_omp = None
_lib_f64 = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


def load_omp() -> C.CDLL:
    """The OpenMP build (all host threads; OMP_NUM_THREADS sets how many)."""
    global _lib_omp
    if _lib_omp is None:
        if not os.path.exists(LIB_OMP_PATH):
            build()
        _lib_omp = _configure(C.CDLL(LIB_OMP_PATH))
    return _lib_omp


def load_f64() -> C.CDLL:
    """The FP64 build: the same restatement with float state widened to double (the `wide` oracles)."""
    global _lib_f64
    if _lib_f64 is None:
        if not os.path.exists(LIB_F64_PATH):
            build()
        _lib_f64 = C.CDLL(LIB_F64_PATH)
        _lib_f64.sbo_network_step.argtypes = [C.POINTER(SboNet64), i32]
        _lib_f64.sbo_contractile_step.argtypes = [C.POINTER(SboNet64), C.POINTER(SboMotors64), i32]
    return _lib_f64


def _configure(lib: C.CDLL) -> C.CDLL:
    lib.sbo_const.restype = C.c_double
    lib.sbo_wang_hash.restype = u32
    lib.sbo_wang_hash.argtypes = [u32]
    lib.sbo_event_bits.restype = u32
    lib.sbo_event_bits.argtypes = [i32, u32, i32, i32, u32, u32]
    lib.sbo_uniform.restype = C.c_float
    lib.sbo_uniform.argtypes = [u32]
    for name in ("sbo_accurate_acos", "sbo_accurate_asin"):
        getattr(lib, name).restype = C.c_double
        getattr(lib, name).argtypes = [C.c_double]
    lib.sbo_angle_from_sin_cos.restype = C.c_double
    lib.sbo_angle_from_sin_cos.argtypes = [C.c_double, C.c_double]
    lib.sbo_rod_drag_si.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double)]
    lib.sbo_sphere_drag_si.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
    lib.sbo_network_step.argtypes = [C.POINTER(SboNet), i32]
    lib.sbo_brownian_bits.argtypes = [i32, u32, i32, i32, u32, C.POINTER(u32)]
    lib.sbo_philox4x32_10.argtypes = [C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    return lib


def P(a: np.ndarray | None) -> C.c_void_p:
    if a is None:
        return C.c_void_p(0)
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def call(name: str, *args):
    """Call an oracle function with numpy arrays (passed as pointers) and python scalars."""
    lib = load()
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(P(a))
        elif a is None:
            conv.append(C.c_void_p(0))
        elif isinstance(a, float):
            conv.append(C.c_float(a))
        elif isinstance(a, (np.float32,)):
            conv.append(C.c_float(float(a)))
        elif isinstance(a, (int, np.integer)):
            conv.append(C.c_int(int(a)) if -2**31 <= int(a) < 2**31 else C.c_uint32(int(a) & 0xffffffff))
        else:
            conv.append(a)
    return getattr(lib, name)(*conv)


class OracleNet:
    """sbo_network_step over host copies of a Scene (DenseContractileHarness.cpuStep filament block)."""

    def __init__(self, scene, lib: C.CDLL | None = None, wide: bool = False):
        fs, xl, p = scene.fs.copy(), (scene.xl.copy() if scene.xl is not None else None), scene.p
        if wide:  # the FP64 run: float64 copies of every float array, the FP64 build of the oracle
            _widen_store(fs)
            if xl is not None:
                _widen_store(xl)
        self.fs, self.xl, self.p, self.wide = fs, xl, p, wide
        self.fdt = np.float64 if wide else np.float32
        n = fs.n
        s = SboNet64() if wide else SboNet()
        s.nRod = n
        for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam",
                  "forceSum", "torqueSum", "randForce", "randTorque", "brownTransScale", "brownRotScale",
                  "end1NbrSlot", "end1NbrSide", "end2NbrSlot", "end2NbrSide"):
            setattr(s, k, P(getattr(fs, k)).value)
        for k in ("params", "chainParams", "xlParams", "offParams", "torsionParams", "boxParams", "gridParams",
                  "boxLen", "gridDims"):
            _fill(getattr(s, k), p[k])
        for k in ("chainOn", "unbindOn", "confineOn", "gridOn", "periodic", "maxCand", "seed", "rng"):
            setattr(s, k, int(p[k]))
        s.stream = int(p["stream"])
        s.actinRadius = float(np.float32(p["actinRadius"]))
        self.keep = []
        if xl is not None:
            s.nLinks = xl.nLinks
            for k in ("linkFilA", "linkFilB", "linkState", "strainPlace", "linkOrientSame", "torqueMagPlace",
                      "activeLinkCount", "loc1", "loc2", "strainHist", "torqueMagHist", "xlinkData"):
                setattr(s, k, P(getattr(xl, k)).value)
            self.csr = {k: np.zeros(n + 1 if "Offsets" in k else (n if "Count" in k else max(1, xl.nLinks)), np.int32)
                        for k in ("segCountA", "segOffsetsA", "segIdxA", "segCountB", "segOffsetsB", "segIdxB")}
            for k, a in self.csr.items():
                setattr(s, k, P(a).value)
        if p["gridOn"]:
            cells = int(p["gridDims"][3])
            mc = int(p["maxCand"])
            self.grid = dict(center=np.zeros(3 * n, self.fdt), boundingRadius=np.zeros(n, self.fdt),
                             bodyCell=np.zeros(n, np.int32), cellCount=np.zeros(cells, np.int32),
                             gridCellOffsets=np.zeros(cells + 1, np.int32), gridCellContents=np.zeros(n, np.int32),
                             candPartner=np.full(n * mc, -1, np.int32), candCount=np.zeros(n, np.int32))
            for k, a in self.grid.items():
                setattr(s, k, P(a).value)
        self.turn = None
        if getattr(scene, "turn", None) is not None:
            self.turn = OracleTurnover(fs, scene.turn, wide)
            s.turn = C.cast(C.pointer(self.turn.s), C.c_void_p).value
        self.form = None
        if p.get("formOn"):
            self.form = OracleFormation(fs, xl, p, wide)
            s.form = C.cast(C.pointer(self.form.s), C.c_void_p).value
            s.formInt = int(p["formInt"])
        self.s = s
        assert not (wide and lib is not None), "the FP64 run uses the FP64 build"
        self.lib = load_f64() if wide else (lib if lib is not None else load())

    def step(self, step0: int, n: int = 1) -> None:
        for t in range(step0, step0 + n):
            self.lib.sbo_network_step(C.byref(self.s), t)

    def cadence(self, step: int) -> None:
        """One turnover cadence alone (sbo_turnover_cadence) on the store, no mechanics step."""
        self.lib.sbo_turnover_cadence(C.byref(self.turn.s), C.byref(self.s), int(step))


class SboMotors(C.Structure):
    _fields_ = (
        [("nM", i32)]
        + [(k, vp) for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam",
                             "forceSum", "torqueSum", "randForce", "randTorque", "brownTransScale", "brownRotScale",
                             "head", "hUVec", "rodUVec", "reach", "boundSeg", "cooldown", "stats", "capStats",
                             "nucleotideState", "forceDotPlace", "bindArc", "forceDotFil", "forceMag", "forceDotHist")]
        + [("bodyParams", C.c_float * 2), ("jointParams", C.c_float * 11), ("kinParams", C.c_float * 14),
           ("nucParams", C.c_float * 8), ("xbParams", C.c_float * 6)]
        + [(k, vp) for k in ("reachSeg", "reachCount", "bondData", "segMotorCount", "segMotorOffsets", "segMotorMyo")]
        + [("nD", i32)]
        + [(k, vp) for k in ("motorA", "motorB", "parallel")]
        + [("dimerParams", C.c_float * 5), ("nBb", i32), ("nMD", i32)]
        + [(k, vp) for k in ("bbCoord", "bbUVec", "bbYVec", "bbZVec", "bbEnd1", "bbEnd2", "bbSegLength", "bbTransGam",
                             "bbRotGam", "bbForceSum", "bbTorqueSum", "bbRandForce", "bbRandTorque", "bbBts", "bbBrs",
                             "bbInvDragY", "headBackboneSlot", "tetherMotor", "attachAxial", "miniData",
                             "bbDimerCount", "bbDimerOffsets", "bbDimerList")]
        + [("miniParams", C.c_float * 3), ("bbBodyParams", C.c_float * 2), ("seedMotor", i32),
           ("seedBackbone", i32), ("ownerStore", vp), ("ownerSlot", vp), ("radial", vp), ("attachCoeffK", vp),
           ("nodeMode", i32)]
    )


class SboMotors64(C.Structure):
    _fields_ = _widen(SboMotors._fields_)


class OracleContractile(OracleNet):
    """sbo_contractile_step over host copies of a ContractileScene (DenseContractileHarness.cpuStep)."""

    def __init__(self, scene, lib: C.CDLL | None = None, wide: bool = False):
        super().__init__(scene, lib, wide)
        mot, dim, mini = scene.mot.copy(), scene.dim.copy(), scene.mini.copy()
        if wide:
            for st in (mot, dim, mini):
                _widen_store(st)
        self.mot, self.dim, self.mini = mot, dim, mini
        n, nM = self.fs.n, mot.nM
        cap = n + nM
        cells = int(self.p["gridDims"][3])
        mc = int(self.p["maxCand"])
        # grid arrays cover filaments + motor heads
        self.grid = dict(center=np.zeros(3 * cap, self.fdt), boundingRadius=np.zeros(cap, self.fdt),
                         bodyCell=np.zeros(cap, np.int32), cellCount=np.zeros(cells, np.int32),
                         gridCellOffsets=np.zeros(cells + 1, np.int32), gridCellContents=np.zeros(cap, np.int32),
                         candPartner=np.full(1, -1, np.int32), candCount=np.zeros(1, np.int32))
        for k, a in self.grid.items():
            setattr(self.s, k, P(a).value)
        m = SboMotors64() if wide else SboMotors()
        m.nM = nM
        body = mot.body
        for k in ("coord", "uVec", "yVec", "zVec", "end1", "end2", "segLength", "bTransGam", "bRotGam", "forceSum",
                  "torqueSum", "randForce", "randTorque", "brownTransScale", "brownRotScale"):
            setattr(m, k, P(getattr(body, k)).value)
        m.hUVec = P(mot.uVec).value
        for k in ("head", "rodUVec", "reach", "boundSeg", "cooldown", "stats", "capStats", "nucleotideState",
                  "forceDotPlace", "bindArc", "forceDotFil", "forceMag", "forceDotHist", "bondData"):
            setattr(m, k, P(getattr(mot, k)).value)
        for k in ("bodyParams", "jointParams", "kinParams", "nucParams", "xbParams"):
            _fill(getattr(m, k), getattr(mot, k))
        self.mreach = dict(reachSeg=np.full(nM * mc, -1, np.int32), reachCount=np.zeros(nM, np.int32),
                           segMotorCount=np.zeros(n, np.int32), segMotorOffsets=np.zeros(n + 1, np.int32),
                           segMotorMyo=np.zeros(max(1, nM), np.int32), ownerStore=np.zeros(cap, np.int32),
                           ownerSlot=np.zeros(cap, np.int32), bbDimerCount=np.zeros(mini.nBb, np.int32),
                           bbDimerOffsets=np.zeros(mini.nBb + 1, np.int32),
                           bbDimerList=np.zeros(max(1, mini.nMD), np.int32))
        for k, a in self.mreach.items():
            setattr(m, k, P(a).value)
        m.nD = dim.nD
        for k in ("motorA", "motorB", "parallel"):
            setattr(m, k, P(getattr(dim, k)).value)
        _fill(m.dimerParams, dim.dimerParams)
        m.nBb, m.nMD = mini.nBb, mini.nMD
        bb = mini.backbone
        for src, dst in (("coord", "bbCoord"), ("uVec", "bbUVec"), ("yVec", "bbYVec"), ("zVec", "bbZVec"),
                         ("end1", "bbEnd1"), ("end2", "bbEnd2"), ("segLength", "bbSegLength"),
                         ("bTransGam", "bbTransGam"), ("bRotGam", "bbRotGam"), ("forceSum", "bbForceSum"),
                         ("torqueSum", "bbTorqueSum"), ("randForce", "bbRandForce"), ("randTorque", "bbRandTorque"),
                         ("brownTransScale", "bbBts"), ("brownRotScale", "bbBrs")):
            setattr(m, dst, P(getattr(bb, src)).value)
        m.bbInvDragY = P(mini.bbInvDragY).value
        m.headBackboneSlot = P(mini.headBackboneSlot).value
        m.tetherMotor = P(mini.motorA).value
        m.attachAxial = P(mini.attachAxial).value
        m.miniData = P(mini.miniData).value
        _fill(m.miniParams, mini.miniParams)
        _fill(m.bbBodyParams, mini.bbBodyParams)
        if hasattr(mini, "radial"):        # node shell: nodes in the backbone slots (NodeStore)
            m.radial = P(mini.radial).value
            m.attachCoeffK = P(mini.attachCoeffK).value
            m.nodeMode = int(mini.mode)
        m.seedMotor = int(self.p["seedMotor"])
        m.seedBackbone = int(self.p["seedBackbone"])
        self.m = m
        if not wide:
            self.lib.sbo_contractile_step.argtypes = [C.POINTER(SboNet), C.POINTER(SboMotors), i32]

    def step(self, step0: int, n: int = 1) -> None:
        for t in range(step0, step0 + n):
            self.lib.sbo_contractile_step(C.byref(self.s), C.byref(self.m), t)


class OracleFormation:
    """An sbo_form over a filament store and crosslinker store (host arrays, updated in place by
    sbo_formation): the formation grid, request arrays and allocator scratch it needs."""

    def __init__(self, fs, xl, p, wide: bool = False):
        n, L = fs.n, xl.nLinks
        f = SboForm64() if wide else SboForm()
        f.nSeg, f.C = n, L
        f.reqCap = int(p["formReqCap"])
        self.keep = []
        for k in ("coord", "uVec", "end1", "end2", "segLength", "end2NbrSlot"):
            setattr(f, k, P(getattr(fs, k)).value)
        self.filState = getattr(fs, "filState", None)
        f.filState = P(self.filState).value if self.filState is not None else None
        for k in ("linkFilA", "linkFilB", "linkState", "strainPlace", "linkOrientSame", "torqueMagPlace",
                  "activeLinkCount", "loc1", "loc2", "strainHist", "torqueMagHist"):
            setattr(f, k, P(getattr(xl, k)).value)
        _fill(f.gridParams, p["formGridParams"])
        _fill(f.gridDims, p["formGridDims"])
        f.periodic = int(p["periodic"])
        _fill(f.boxLen, p["boxLen"])
        _fill(f.formParams, p["formParams"])
        f.mode = int(p["formMode"])
        cells = int(p["formGridDims"][3])
        rc = f.reqCap
        sizes = dict(filID=n, filIDScratch=n, bodyCell=n, cellCount=cells, gridCellOffsets=cells + 1,
                     gridCellContents=n, candCountSeg=n, candBaseSeg=n + 1, reqFilA=rc, reqFilB=rc, reqOrient=rc,
                     gatePass=rc, minCand=n, acceptFlag=rc, freeCount=max(1, L), freeOffsets=L + 1,
                     freeList=max(1, L), rankOffsets=rc + 1)
        self.a = {k: np.zeros(v, np.int32) for k, v in sizes.items()}
        fdt = np.float64 if wide else np.float32
        self.a.update(center=np.zeros(3 * n, fdt), reqLoc1=np.zeros(rc, fdt), reqLoc2=np.zeros(rc, fdt))
        for k, arr in self.a.items():
            setattr(f, k, P(arr).value)
        self.s = f

    @property
    def n_candidates(self) -> int:
        return int(self.s.formCounts[0])


class OracleTurnover:
    """An sbo_turnover over a filament store (host arrays, updated in place by sbo_turnover_cadence): a copy
    of the scene's TurnoverStore, the cadence scratch, the pool and the last cadence's counters."""

    def __init__(self, fs, ts, wide: bool = False):
        t = ts.copy()
        if wide:
            _widen_store(t)
        self.ts = t
        n = t.C
        f = SboTurnover64() if wide else SboTurnover()
        f.C = n
        fdt = np.float64 if wide else np.float32
        self.a = {k: getattr(t, k) for k in TURN_I + ("nucFrac", "cofFrac")}
        self.a["bTransDiff"] = fs.bTransDiff
        self.a["bRotDiff"] = fs.bRotDiff
        for k in TURN_SCRATCH_I:
            self.a[k] = np.zeros(n, np.int32)
        for k in TURN_SCRATCH_F:
            self.a[k] = np.zeros(3 * n, fdt)
        for k in TURN_OFFS:
            self.a[k] = np.zeros(n + 1, np.int32)
        for k, arr in self.a.items():
            setattr(f, k, P(arr).value)
        f.interval, f.seed, f.growthOn, f.proxy = int(t.interval), int(t.seed), int(t.growthOn), int(t.proxy)
        for k in ("growParams", "splitParams", "dragParams", "depolyParams", "depolyRateParams", "agingParams",
                  "cofilinParams", "birthParams"):
            _fill(getattr(f, k), getattr(t, k))
        f.onRate, f.biochemDeltaT, f.uMPerMonomer = float(t.onRate), float(t.biochemDeltaT), float(t.uMPerMonomer)
        f.conc = float(t.pool[0])
        f.taken, f.returned = int(t.ledger[0]), int(t.ledger[1])
        self.s = f

    @property
    def conc(self) -> float:
        return float(self.s.conc)

    @property
    def ledger(self):
        return int(self.s.taken), int(self.s.returned)

    @property
    def last(self):
        return list(self.s.last)
