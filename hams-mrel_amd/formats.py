"""HAMS on-disk formats around the solve: the control file, the frequency
conventions, and the HAMS- and WAMIT-style coefficient files, so that a HAMS
input directory runs unchanged and its outputs compare line by line.

* ``read_control_file``: ReadControlFile (InputFiles.f90:46-224), the same
  record order; each value is taken from the end of its record, which for the
  reference's fixed-column files is what its edit descriptors read.
* ``wave_properties``: CalWaveProperts (ImplementSubs.f90:40-151), the input
  frequency types 1-5 to (w1, wk, tp, wl, v) and the output value OUFR.
* ``problem_from_control``: the bodies of ReadPanelMeshMulti
  (ReadPanelMeshMulti.f90:46-111) placed by the control file's LCS and
  reference centres, with the reference's constants; with If_remove_irr_freq
  = 1 the interior waterplane meshes too (WaterPlaneMesh.pnl for one body,
  WaterPlaneMesh_k.pnl per body where present, HAMS_Prog.f90:169-173,
  :399-418).
* ``frequency_records``: the frequency loop's records in the reference's
  order, the zero- and infinite-frequency limits first when
  0_inf_frequency_limits = 1 (InputFiles.f90:88-110, ImplementSubs.f90:45-70).
* ``run_input_directory``: an input directory solved on the GPU record by
  record and written as the HAMS and WAMIT files.
* ``write_hams``: the OAMASS/ODAMPING/OEXFOR files (PotentWavForceMulti.f90:
  119, :202-205; FORMAT 1000 F7.3,1x,F7.3,1x,6E14.5 and 1010 ... 6E14.6).
* ``write_wamit_1`` / ``write_wamit_3``: the WAMIT-style .1 and .3 files
  (PotentWavForceMulti.f90:122, :209-216; FORMAT 1020 ES14.6,2I6,2ES14.6 and
  1030 2ES14.6,I6,4ES14.6), with ReorderAmssDamp's row-major order
  (PotentWavForceMulti.f90:235) as an option. The .3 phase keeps the
  reference's ATAN2D(...)*180/pi (degrees converted twice,
  PotentWavForceMulti.f90:117) unless ``phase_degrees_twice=False``.
"""
import math
import os
from dataclasses import dataclass, field

import numpy as np

import meshes

G_MODULE = 9.80665     # WavDynMods.f90:146
G_DISPERSION = 9.807   # WavDynSubs.f90:58


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


@dataclass
class ControlFile:
    h: float
    sybo: int
    inft: int
    ouft: int
    values: np.ndarray              # the input frequency values (type inft)
    nbody: int
    lcs: np.ndarray                 # nbody x 4: x, y, z, rotation about z [deg]
    headings_deg: np.ndarray
    xr: np.ndarray                  # nbody x 3 reference centres
    refl: float
    isol: int
    irsp: int
    nthread: int
    isolv: int = 1
    xfp: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    separate_wamit: bool = False


class _Records:
    def __init__(self, path):
        with open(path) as f:
            self.lines = f.read().split("\n")
        self.i = 0

    def next(self):
        if self.i >= len(self.lines):
            raise ValueError("ControlFile.in ends early (record %d)" % (self.i + 1))
        s = self.lines[self.i]
        self.i += 1
        return s

    def skip(self, k):
        for _ in range(k):
            self.next()

    def tail(self, k=1):
        tok = self.next().split()
        if len(tok) < k:
            raise ValueError("ControlFile.in record %d: %d values expected" % (self.i, k))
        return tok[-k:]

    def values(self, count):
        # a list-directed read: values across as many records as it takes
        out = []
        while len(out) < count:
            out += [_num(t) for t in self.next().replace(",", " ").split()]
        return np.array(out[:count])


def read_control_file(path):
    """ControlFile.in (InputFiles.f90:46-224); ``path`` is the file or its directory"""
    if os.path.isdir(path):
        path = os.path.join(path, "ControlFile.in")
    if not os.path.exists(path):
        raise FileNotFoundError("Error opening ControlFile.in for reading: " + path)
    r = _Records(path)
    r.skip(2)
    h = _num(r.tail()[0])
    r.skip(2)
    sybo = int(r.tail()[0])
    inft = int(r.tail()[0])
    ouft = int(r.tail()[0])
    npet = int(r.tail()[0])
    if sybo not in (0, 1):
        sybo = 0        # the reference warns and takes 0
    if npet > 0:
        vals = r.values(npet)
    elif npet < 0:
        wk1, dwk = _num(r.tail()[0]), _num(r.tail()[0])
        vals = wk1 + np.arange(-npet) * dwk
    else:
        vals = np.zeros(0)
    r.skip(2)
    body_check = r.next()
    nbody = int(r.tail()[0]) if "multi" in body_check else 1
    if nbody < 1:
        raise ValueError("The number of bodies must be greater than or equal to 1.")
    lcs = np.zeros((nbody, 4))
    if nbody > 1:
        for b in range(nbody):
            lcs[b] = [_num(t) for t in r.tail(4)]
        r.skip(3)
    nbeta = int(r.tail()[0])
    if nbeta > 0:
        hd = r.values(nbeta)
    elif nbeta < 0:
        b1, db = _num(r.tail()[0]), _num(r.tail()[0])
        hd = b1 + np.arange(-nbeta) * db
    else:
        hd = np.zeros(0)
    r.skip(2)
    xr = np.array([[_num(t) for t in r.tail(3)] for _ in range(nbody)])
    refl = _num(r.tail()[0])
    isol = int(r.tail()[0])
    irsp = int(r.tail()[0])
    nthread = int(r.tail()[0])
    isolv = 1
    line = r.next()
    if "Solver_type" in line:
        try:
            isolv = int(line.split()[-1])
        except ValueError:
            isolv = 1
        if isolv not in (1, 2):
            isolv = 1
    else:
        r.i -= 1
    r.skip(2)
    nfp = int(r.tail()[0])
    xfp = np.array([r.values(3) for _ in range(nfp)]).reshape(nfp, 3)
    separate = False
    try:
        r.skip(2)
        line = r.next()
        if "separate" in line or "Separate" in line:
            separate = int(line.split()[-1]) != 0
    except ValueError:
        pass
    if nbody == 1:
        separate = False
    return ControlFile(h=h, sybo=sybo, inft=inft, ouft=ouft, values=np.asarray(vals, dtype=np.float64),
                       nbody=nbody, lcs=lcs, headings_deg=np.asarray(hd, dtype=np.float64), xr=xr, refl=refl,
                       isol=isol, irsp=irsp, nthread=nthread, isolv=isolv, xfp=xfp, separate_wamit=separate)


def _wave_number(w1, h, g_disp):
    import hm_bem
    return hm_bem.dispersion(w1, h, g_disp, nrt=1)[0]


def wave_properties(value, inft, ouft, h, g=G_MODULE, g_disp=G_DISPERSION):
    """CalWaveProperts (ImplementSubs.f90:71-149) for one input value:
    dict(w1, wk, tp, wl, v, oufr)"""
    deep = h <= 0.0
    if inft == 1:                       # deep-water wave number
        v = value
        w1 = math.sqrt(g * v)
        wk = v if deep else _wave_number(w1, h, g_disp)
    elif inft == 2:                     # finite-depth wave number
        wk = value
        v = wk if deep else wk * math.tanh(wk * h)
        w1 = math.sqrt(g * v)
    elif inft in (3, 4):                # wave frequency, wave period
        w1 = value if inft == 3 else 2.0 * math.pi / value
        v = w1 * w1 / g
        wk = v if deep else _wave_number(w1, h, g_disp)
    elif inft == 5:                     # wave length
        wk = 2.0 * math.pi / value
        v = wk if deep else wk * math.tanh(wk * h)
        w1 = math.sqrt(g * v)
    else:
        raise ValueError("Input_frequency_type must be 1..5, got %r" % inft)
    out = dict(w1=w1, wk=wk, tp=2.0 * math.pi / w1, wl=2.0 * math.pi / wk, v=v)
    if ouft not in (1, 2, 3, 4, 5):
        raise ValueError("Output_frequency_type must be 1..5, got %r" % ouft)
    out["oufr"] = out[("v", "wk", "w1", "tp", "wl")[ouft - 1]]
    return out


def _waterplane(path, lcs, log):
    """interior waterplane mesh of one body (ReadWTPLMesh, ReadPanelMesh.f90:74-
    98: every node on z = 0, else the reference stops)"""
    wp = meshes.read_pnl(path, lcs=lcs)
    bad = np.nonzero(np.abs(wp.xyz[:, 2]) > 1e-10)[0]
    if len(bad):
        raise ValueError("Error: Z Coordinate is not zero at Node No. %d of %s" % (bad[0] + 1, path))
    return wp


def problem_from_control(inp, hull="HullMesh_{k}.pnl", single_hull="HullMesh.pnl", log=print):
    """meshes.Problem of a HAMS input directory: the bodies' .pnl meshes placed by
    the control file's LCS and reference centres, the frequencies converted by
    wave_properties, the headings in radians, the reference's constants.
    If_remove_irr_freq = 1 adds the interior waterplane meshes (one body:
    WaterPlaneMesh.pnl, which must exist; several: WaterPlaneMesh_k.pnl for the
    bodies that have one, HAMS_Prog.f90:399-418). Wave_diffrac_solution must be
    1 (total potential); 2 (scattered potential only) is not provided by the
    C-ABI and raises NotImplementedError. The frequency list holds the regular
    frequencies only; frequency_records adds the zero- and infinite-frequency
    limits of 0_inf_frequency_limits = 1."""
    cf = read_control_file(inp)
    if cf.isol not in (1, 2):
        raise ValueError("Error: The input for ISOL should be either 1 or 2.")
    if cf.isol == 2:
        raise NotImplementedError("Wave_diffrac_solution 2 (scattered potential, PotentWavForceMulti.f90:72) "
                                  "is not provided: the excitation of the GPU path includes the incident potential")
    if cf.irsp not in (0, 1):
        raise ValueError("If_remove_irr_freq must be 0 or 1, got %r" % cf.irsp)
    bodies = []
    for b in range(cf.nbody):
        name = single_hull if cf.nbody == 1 else hull.format(k=b + 1)
        lcs = tuple(cf.lcs[b])
        body = meshes.read_pnl(os.path.join(inp, name), lcs=lcs, xr=tuple(cf.xr[b]))
        if cf.irsp == 1:
            wname = "WaterPlaneMesh.pnl" if cf.nbody == 1 else "WaterPlaneMesh_%d.pnl" % (b + 1)
            wpath = os.path.join(inp, wname)
            if os.path.exists(wpath) or cf.nbody == 1:
                wp = _waterplane(wpath, lcs, log)
                body.ixyz, body.incn, body.incon, body.ilocal = wp.xyz, wp.ncn, wp.ncon, wp.local
            else:
                log(" Warning: The waterplane mesh file does not exist for Body %d" % (b + 1))
        bodies.append(body)
    w = [wave_properties(v, cf.inft, cf.ouft, cf.h)["w1"] for v in cf.values]
    p = meshes.Problem(bodies, h=cf.h, omegas=np.array(w), beta=np.deg2rad(cf.headings_deg),
                       name=os.path.basename(os.path.normpath(inp)))
    apply_generalized_modes(p, inp)
    return meshes.reference_parity_constants(p), cf


def frequency_records(cf, g=G_MODULE, g_disp=G_DISPERSION):
    """The records of the reference's frequency loop (NPER of them, InputFiles.f90:
    88-110): with 0_inf_frequency_limits = 1 the zero-frequency record (V = WK =
    W1 = 0, TP = WL = -1) and the infinite-frequency record (V = WK = W1 = -1,
    TP = WL = 0) come first (ImplementSubs.f90:45-70), then the listed values
    through wave_properties. Each record: dict(kind 'zero' | 'inf' | 'wave',
    w1, wk, tp, wl, v, oufr)."""
    out = []
    if cf.sybo == 1:
        for kind, val, per in (("zero", 0.0, -1.0), ("inf", -1.0, 0.0)):
            r = dict(kind=kind, w1=val, wk=val, tp=per, wl=per, v=val)
            r["oufr"] = r[("v", "wk", "w1", "tp", "wl")[cf.ouft - 1]]
            out.append(r)
    for v in cf.values:
        r = wave_properties(v, cf.inft, cf.ouft, cf.h, g, g_disp)
        r["kind"] = "wave"
        out.append(r)
    return out


def run_input_directory(inp, outdir=None, precision=None, ctx=None):
    """A HAMS input directory solved on the GPU in the reference's record order
    (HAMS_Prog.f90:271-300, :629-724): each record is a frequency solve or, for
    0_inf_frequency_limits = 1, a zero- or infinite-frequency limit solve; with
    outdir the HAMS files (outdir/Hams_format) and the WAMIT .1/.3 files
    (outdir/Wamit_format, the limits' records as the reference writes them)
    are written. Returns (records, amas, bdmp, exfc) with amas, bdmp
    (nrec, 6 nbody, 6 nbody) and exfc (nrec, 6 nbody, nbeta)."""
    import hm_bem
    if precision is None:
        precision = hm_bem.HM_FP64
    p, cf = problem_from_control(inp)
    recs = frequency_records(cf, p.g, p.g_disp)
    own = ctx is None
    if own:
        ctx = hm_bem.Context()
    try:
        ctx.setup(p)
        res = []
        for r in recs:
            if r["kind"] == "wave":
                res.append(ctx.solve_frequency(r["w1"], precision))
            else:
                res.append(ctx.solve_limit(0 if r["kind"] == "zero" else 1, precision))
    finally:
        if own:
            ctx.close()
    amas = np.array([x["amas"] for x in res])
    bdmp = np.array([x["bdmp"] for x in res])
    exfc = np.array([x["exfc"] for x in res])
    if outdir is not None:
        wk = [r["wk"] for r in recs]
        w1 = [r["w1"] for r in recs]
        write_hams(os.path.join(outdir, "Hams_format"), wk, w1, amas, bdmp, exfc)
        wam = os.path.join(outdir, "Wamit_format")
        os.makedirs(wam, exist_ok=True)
        oufr = [r["oufr"] for r in recs]
        limit = [r["kind"] != "wave" for r in recs]
        write_wamit_1(os.path.join(wam, "AmssDamp.1"), oufr, w1, amas, bdmp, p.rho, limit=limit)
        write_wamit_3(os.path.join(wam, "ExcForce.3"), oufr, cf.headings_deg, exfc, p.rho, p.g, p.amp, limit=limit)
    return recs, amas, bdmp, exfc


def _list_directed(recs, start, count):
    """count values of one list-directed READ beginning at record `start`: blank
    records are skipped, values run across records, and the rest of the record
    holding the last value is dropped. Returns (tokens, next record); fewer
    tokens than count at the end of the file."""
    out, r = [], start
    while len(out) < count and r < len(recs):
        tok = recs[r].replace(",", " ").split()
        r += 1
        out += tok[:count - len(out)]
    return out, r


def apply_generalized_modes(p, inp, log=print):
    """The generalized-mode override of CalNormalsMulti (ReadPanelMeshMulti.f90:
    166-203): gen_mod_{body}_{dof}.txt (body and dof 1-based) holds a pair
    "panel_id normal_value" per hull panel of the body, each read by one
    list-directed READ(99,*) (:189): blank records are skipped, a pair may run
    across records, and what follows the pair on its record is dropped. Each
    pair replaces that dof's normal component of the panel; a value that does
    not read, the end of the file before every panel has a pair, or a panel
    number out of range (which the reference would write out of bounds) ends
    the file with a warning, as in the reference. Returns the (body, dof) pairs
    overridden."""
    done = []
    for b in range(p.nbody):
        n0, nb = int(p.body_start[b]), int(p.body_start[b + 1] - p.body_start[b])
        for dof in range(1, 7):
            path = os.path.join(inp, "gen_mod_%d_%d.txt" % (b + 1, dof))
            if not os.path.exists(path):
                continue
            with open(path) as fh:
                recs = fh.read().split("\n")
            r = 0
            for k in range(nb):
                tok, r = _list_directed(recs, r, 2)
                try:
                    pid, val = int(tok[0]), _num(tok[1])
                    if not 1 <= pid <= nb:
                        raise ValueError
                except (IndexError, ValueError):
                    log("  Warning: Could not read line %d from %s" % (k + 1, path))
                    break
                p.dxyz_p[n0 + pid - 1, dof - 1] = val
                if getattr(p, "frames", False):
                    p.local["dxyz_p"][n0 + pid - 1, dof - 1] = val
            done.append((b + 1, dof))
    return done


# ---- Fortran edit descriptors ------------------------------------------------------------------
def fortran_e(x, w, d):
    """Ew.d: 0.ddddE+ee right-justified in w"""
    x = float(x)
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, e = ("%.*e" % (d - 1, abs(x))).split("e")
        s = ("-" if x < 0 else "") + "0." + m.replace(".", "") + "E%+03d" % (int(e) + 1)
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def fortran_es(x, w, d):
    """ESw.d: d.ddddE+ee right-justified in w"""
    s = "%.*E" % (d, float(x))
    return "*" * w if len(s) > w else s.rjust(w)


def fortran_f(x, w, d):
    s = "%.*f" % (d, float(x))
    return "*" * w if len(s) > w else s.rjust(w)


# ---- writers -----------------------------------------------------------------------------------
def write_hams(outdir, wk, w1, amas, bdmp, exfc):
    """OAMASS1-6, ODAMPING1-6, OEXFOR1-6 (Hams_format). wk, w1: (nfreq,);
    amas, bdmp: (nfreq, 6 nbody, 6 nbody); exfc: (nfreq, 6 nbody, nbeta).
    Per frequency: radiation lines body pair by body pair (row body outer),
    one line of the row's 6 columns per file; excitation lines heading by
    heading, body by body."""
    amas, bdmp, exfc = np.asarray(amas), np.asarray(bdmp), np.asarray(exfc)
    nf, m = amas.shape[0], amas.shape[1]
    nb = m // 6
    os.makedirs(outdir, exist_ok=True)
    for md in range(6):
        la, lb, lf = [], [], []
        for f in range(nf):
            head = fortran_f(wk[f], 7, 3) + " " + fortran_f(w1[f], 7, 3) + " "
            for b1 in range(nb):
                for b2 in range(nb):
                    la.append(head + "".join(fortran_e(v, 14, 5) for v in amas[f, 6 * b1 + md, 6 * b2:6 * b2 + 6]))
                    lb.append(head + "".join(fortran_e(v, 14, 5) for v in bdmp[f, 6 * b1 + md, 6 * b2:6 * b2 + 6]))
            for ih in range(exfc.shape[2]):
                for b1 in range(nb):
                    z = exfc[f, 6 * b1 + md, ih]
                    lf.append(head + fortran_e(z.real, 14, 6) + fortran_e(z.imag, 14, 6))
        for stem, lines in (("OAMASS", la), ("ODAMPING", lb), ("OEXFOR", lf)):
            with open(os.path.join(outdir, "%s%d.txt" % (stem, md + 1)), "w") as fh:
                fh.write("\n".join(lines) + "\n")


def write_wamit_1(path, oufr, w1, amas, bdmp, rho, limit=None, row_major=True):
    """WAMIT .1: OUFR, i, j, A_ij / rho, B_ij / (rho w1). limit[f] true marks a
    zero- or infinite-frequency record (period -1 or 0), written without B.
    row_major: ReorderAmssDamp's order (i, then j) within each frequency;
    otherwise the order the reference writes, body pair by body pair."""
    amas, bdmp = np.asarray(amas), np.asarray(bdmp)
    nf, m = amas.shape[0], amas.shape[1]
    nb = m // 6
    out = []
    for f in range(nf):
        if row_major:
            order = [(i, j) for i in range(m) for j in range(m)]
        else:
            order = [(6 * b1 + i, 6 * b2 + j) for b1 in range(nb) for b2 in range(nb)
                     for i in range(6) for j in range(6)]
        lim = bool(limit[f]) if limit is not None else False
        for i, j in order:
            s = fortran_es(oufr[f], 14, 6) + "%6d%6d" % (i + 1, j + 1) + fortran_es(amas[f, i, j] / rho, 14, 6)
            if not lim:
                s += fortran_es(bdmp[f, i, j] / (rho * w1[f]), 14, 6)
            out.append(s)
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def write_wamit_3(path, oufr, headings_deg, exfc, rho, g, amp=1.0, phase_degrees_twice=True, limit=None):
    """WAMIT .3: OUFR, heading [deg], i, |F|, phase, Re, Im of the excitation
    normalised by rho g A, with Re, Im the reference's (-Im F, -Re F) / (rho g A)
    and the phase ATAN2D(Im, Re) (times 180/pi again by default, as the reference
    writes it). Records per frequency: heading by heading, body by body.
    limit[f] true marks a zero- or infinite-frequency record (period -1 or 0),
    for which the reference writes no excitation (PotentWavForceMulti.f90:120)."""
    exfc = np.asarray(exfc)
    nf, m, nbeta = exfc.shape
    nb = m // 6
    out = []
    for f in range(nf):
        if limit is not None and limit[f]:
            continue
        for ih in range(nbeta):
            for b in range(nb):
                for md in range(6):
                    z = exfc[f, 6 * b + md, ih] / (rho * g * amp)
                    nrel, nimg = -z.imag, -z.real
                    phs = math.degrees(math.atan2(nimg, nrel))
                    if phase_degrees_twice:
                        phs = phs * 180.0 / math.pi
                    out.append(fortran_es(oufr[f], 14, 6) + fortran_es(headings_deg[ih], 14, 6)
                               + "%6d" % (6 * b + md + 1)
                               + "".join(fortran_es(v, 14, 6) for v in (abs(z), phs, nrel, nimg)))
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
