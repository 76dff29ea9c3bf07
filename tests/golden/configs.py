"""Synthetic starting configurations of the fixtures and tests (data only)."""
import numpy as np


def lattice_start(n, L, seed):
    """Simple-cubic lattice, m = ceil(n^(1/3)) per edge, sites chosen by a
    seeded permutation, small seeded jitter, Shoemake-uniform quaternions."""
    rng = np.random.default_rng(seed)
    m = int(np.ceil(round(n ** (1.0 / 3.0), 9)))
    while m ** 3 < n:
        m += 1
    g = (np.arange(m) + 0.5) * L / m
    pts = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    idx = np.sort(rng.permutation(len(pts))[:n])
    com = pts[idx] + rng.normal(0.0, 0.05, (n, 3))
    u = rng.random((n, 3))
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]),
                  np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]),
                  np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], 1)
    return com, q


def mixture_deck(com, q, L, rc, taper, dt, nsteps, seed, nose=None, model="TIP4P", nlog=1,
                 snpo=0, snpv=0, fast=None):
    """A Mixture input deck (the reference's tags and payload layout,
    Mixture.F90:597-1140): @DTPS, @MDLP, @NLOG, @RCOA, @BXLA, @VOXA, @VSET,
    optional @SNPO / @SNPV (snapshot intervals), optional @FAST (custom
    kernels), optional @NOS2 (T, q, dump), @ID08, @NX4A with x y z a b c d
    per line."""
    n = len(com)
    dt, rc, taper, L = float(dt), float(rc), float(taper), float(L)
    deck = ["@DTPS", repr(dt), "@MDLP", str(int(nsteps)), "@NLOG", str(int(nlog)), "@RCOA",
            "%r %r" % (rc, taper), "@BXLA", repr(L), "@VOXA", "%r %r %r" % (rc, rc, rc),
            "@VSET", "300.0 %d" % int(seed)]
    if snpo:
        deck += ["@SNPO", str(int(snpo))]
    if snpv:
        deck += ["@SNPV", str(int(snpv))]
    if fast is not None:
        deck += ["@FAST", "1" if fast else "0"]
    if nose is not None:
        deck += ["@NOS2", "%r %r %r 0 0 0 0 0 0" % tuple(float(v) for v in nose)]
    deck += ["@ID08", model, "@NX4A", str(n)]
    deck += [" ".join("%.17e" % v for v in list(com[i]) + list(q[i])) for i in range(n)]
    return "\n".join(deck) + "\n"


def parse_logd(text):
    """The @LOGD rows of a Mixture-format log as an array (one row per line
    after each @LOGD tag)."""
    lines = text.split("\n")
    rows = [lines[i + 1] for i, l in enumerate(lines[:-1]) if l.strip() == "@LOGD"]
    return np.array([[float(v) for v in r.split()] for r in rows])


def parse_snapshots(text):
    """The snapshot blocks of a Mixture-format log, in order: for each
    @ASTP block a dict with step, box (3), ncmp, fixc (or None), id, kind
    ("NX4A" or "WTG3") and rows (n x 7 or n x 13)."""
    toks = text.split("\n")
    out, i = [], 0
    while i < len(toks):
        if toks[i].strip() != "@ASTP":
            i += 1
            continue
        b = {"step": int(toks[i + 1]), "fixc": None}
        i += 2
        while True:
            tag = toks[i].strip()
            if tag == "@BOX3":
                b["box"] = [float(v) for v in toks[i + 1].split()]
                i += 2
            elif tag == "@NCMP":
                b["ncmp"] = int(toks[i + 1])
                i += 2
            elif tag == "@FIXC":
                b["fixc"] = toks[i + 1].strip()
                i += 2
            elif tag == "@ID08":
                b["id"] = toks[i + 1][:8]
                i += 2
            elif tag in ("@NX4A", "@WTG3"):
                b["kind"] = tag[1:]
                n = int(toks[i + 1])
                i += 2
                width = 7 if tag == "@NX4A" else 13
                vals = []
                while len(vals) < n * width:
                    vals += [float(v) for v in toks[i].split()]
                    i += 1
                b["rows"] = np.array(vals).reshape(n, width)
                break
            else:
                raise ValueError("unexpected line in a snapshot block: %r" % toks[i])
        out.append(b)
    return out


def tag_decks(g):
    """Variants of the 20-step Nose NVT deck of md_tip4p_216_nvt.npz that use
    the other forms of the reference's tags (Mixture.F90:712-717 hands each
    tag to its module's loader): @NOSE (NoseModule.F90:90-103), @VOXN
    (GridModule.F90:139-159), @BOX3 (BoxModule.F90:109-116) and @BOOK
    (BookModule.F90:48-64).  Returns {name: deck text}."""
    L = float(g["L"])
    base = mixture_deck(g["com"], g["quat"], L, float(g["rc"]), float(g["taper"]),
                        float(g["dt"]), int(g["nsteps"]), int(g["seed"]), tuple(g["nose"]))
    lines = base.split("\n")

    def swap(tag, new_tag, payload):
        out = list(lines)
        k = out.index(tag)
        out[k], out[k + 1] = new_tag, payload
        return "\n".join(out)

    T, Q, dump = (float(v) for v in g["nose"])
    return {
        # the older thermostat tag: no zetasum, the mass in the older unit
        "nose": swap("@NOS2", "@NOSE", "%r 1500.0 %r 0 0 0 0 0" % (T, dump)),
        # a fixed grid of one cell: the reference's fixed-grid stencil then
        # covers every pair within rc (with more cells of this small box it
        # does not, and its interacting set is not the full one)
        "voxn": swap("@VOXA", "@VOXN", "1 1 1"),
        "box3": swap("@BXLA", "@BOX3", "%r %r %r" % (L, L, L)),
        # a booked list rebuilt every 10 steps with a 1 A margin
        "book": base.replace("@ID08", "@BOOK\n10 1.0\n@ID08", 1),
    }


def mc2_files(com, q, L, rc, taper, seed, temp, sweeps, umbrellas=(), refhist=None):
    """The input of the reference's mc2 for one node (mc2util.F90:119-300):
    (config file text, node input text).  @MCLP = @INNR = @SNPO = sweeps,
    @NLOG 1, no grid and no bookkeeping (mc2 then uses the all-pairs list).
    umbrellas: (kind, center, coeffi) per @GEUM (mc2util.F90:174-191), after
    an @HIST of refhist[21] (bins of 0.1 from -1) when refhist is given."""
    n = len(com)
    deck = ["@SEED", str(seed), "@MCLP", str(sweeps), "@INNR", str(sweeps), "@NLOG", "1",
            "@SNPO", str(sweeps), "@TEMP", repr(temp), "@RCOA", "%r %r" % (rc, taper),
            "@BXLA", repr(L)]
    if refhist is not None:
        deck += ["@HIST", "21 0.1 -1.0"] + ["%r" % float(v) for v in refhist]
    for kind, center, coeffi in umbrellas:
        deck += ["@GEUM", kind, "%r %r" % (float(center), float(coeffi))]
    deck += ["@NCMP", "1", "@ID08", "TIP4P", "@NX4A", str(n)]
    deck += [" ".join("%.17e" % v for v in list(com[i]) + list(q[i])) for i in range(n)]
    return "@MPIN\n1\nnode0.in\n", "\n".join(deck) + "\n"


def parse_mc2o_rows(text):
    """The @MC2O lines of an mc2 output file: loop, energy, 1/beta, then
    (op, center, coeffi) per umbrella, as float rows."""
    lines = text.split("\n")
    return np.array([[float(v) for v in lines[i + 1].split()] for i, l in enumerate(lines)
                     if l.strip() == "@MC2O"])


def parse_mc2_out(text, n):
    """(energies of the @MC2O lines, delta, rtratio, final com, final quat)
    of an mc2 output file."""
    lines = text.split("\n")
    eps = [float(lines[i + 1].split()[1]) for i, l in enumerate(lines) if l.strip() == "@MC2O"]
    k = lines.index("@DELT")
    assert int(lines[k + 1]) == n
    dr = np.array([[float(v) for v in lines[k + 2 + i].split()] for i in range(n)])
    k = max(i for i, l in enumerate(lines) if l.strip() == "@NX4A")
    assert int(lines[k + 1]) == n
    cq = np.array([[float(v) for v in lines[k + 2 + i].split()] for i in range(n)])
    return np.array(eps), dr[:, 0], dr[:, 1], cq[:, :3], cq[:, 3:]
