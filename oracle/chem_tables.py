"""Oracle-side restatement of COREFL's Chemkin readers (TEST INFRASTRUCTURE ONLY).

Builds the species / reaction / transport tables the per-point physics reads,
following the reference's readers:

* mechanism ELEMENTS / SPECIES blocks: src/ChemData.cpp:11-76
* NASA-7 therm.dat (Combustion2Part build): src/ChemData.cpp:545-862
* NASA-9 therm.dat (multi-range, molecular weight from the file): src/ChemData.cpp:865-1034
* two-temperature metadata (theta_v, electronic levels, Millikan-White pair
  coefficients, Park cross-section): src/ChemData.cpp:130-232
* AIR-5 control-temperature hook: src/ChemData.cpp:1248-1262, 1294-1297
* molecular weight summed over elements in std::map (alphabetical) order with
  the IUPAC standard atomic weights: src/ChemData.cpp:706-711, src/Element.cpp:13-30
* transport file + derived pair tables: src/ChemData.cpp:1037-1162
* reactions incl. third-body, LOW/TROE fall-off, DUPLICATE merge, REV:
  src/ChemData.cpp:1165-1493
* device tables imw = 1/mw, R_l = R_u/mw: src/DParameter.cu:80-88

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
"""
from __future__ import annotations

import math
import numpy as np

R_U = 8314.462618          # J/(kmol K), src/Constants.h:19
R_C = 1.98720425864083     # cal/(mol K), src/Constants.h:21
BOLTZMANN = 1.380649e-23   # J/K, src/Constants.h:31
PI = 3.141592653589793

# IUPAC standard atomic weights for the elements the shipped mechanisms use.
ATOMIC_WEIGHT = {"H": 1.008, "HE": 4.002602, "C": 12.011, "N": 14.007, "O": 15.999, "AR": 39.95}


def read_species(mech_path):
    """Element and species lists in file order (src/ChemData.cpp:11-76)."""
    toks = []
    with open(mech_path) as f:
        for line in f:
            s = line.split("!")[0]
            toks.extend(s.split())
    it = iter(toks)
    for t in it:
        if t in ("ELEMENTS", "ELEM"):
            break
    elems = []
    for t in it:
        u = t.upper()
        if u == "END":
            continue
        if u in ("SPECIES", "SPEC"):
            break
        elems.append(u)
    species = []
    for t in it:
        u = t.upper()
        if u == "END":
            continue
        if u in ("REACTIONS", "REAC", "THERMO"):
            break
        species.append(u)
    return elems, species


def read_therm_nasa7(therm_path, species, elems):
    """NASA-7 reader, Combustion2Part branch (src/ChemData.cpp:545-862)."""
    ns = len(species)
    idx = {s: i for i, s in enumerate(species)}
    lines = open(therm_path).read().splitlines()
    pos = 0
    while pos < len(lines) and not lines[pos].upper().startswith("THERMO"):
        pos += 1
    pos += 1
    tl = tm = th = None
    while pos < len(lines):
        s = lines[pos]
        pos += 1
        if not s or s[0] == "!":
            continue
        tl, tm, th = (float(v) for v in s.split()[:3])
        break
    t_low = np.full(ns, tl)
    t_mid = np.full(ns, tm)
    t_high = np.full(ns, th)
    low = np.zeros((ns, 7))
    high = np.zeros((ns, 7))
    mw = np.zeros(ns)
    comp = np.zeros((ns, len(elems)), dtype=int)
    have = set()
    n_read = 0
    while pos < len(lines):
        s = lines[pos].upper()
        pos += 1
        if not s or s[0] == "!":
            continue
        key = s.split()[0]
        if key == "END":
            continue
        if key in ("REACTIONS", "REAC"):
            break
        if n_read >= ns:
            break
        name = s[:18].split()[0]
        if name not in idx or name in have:
            pos += 3
            continue
        sp = idx[name]
        t_low[sp] = float(s[45:55])
        t_high[sp] = float(s[55:65])
        tmid_field = s[65:75].split()
        if tmid_field:
            t_mid[sp] = float(tmid_field[0])
        for i in range(4):
            field = s[24 + i * 5: 29 + i * 5].lstrip()
            if not field or field.startswith("0"):
                break
            parts = field.split()
            el = parts[0]
            if el not in elems:
                continue
            comp[sp, elems.index(el)] = int(parts[1]) if len(parts) > 1 else 0
        w = 0.0
        for el in sorted(elems):
            w += ATOMIC_WEIGHT[el] * comp[sp, elems.index(el)]
        mw[sp] = w
        l2, l3, l4 = lines[pos], lines[pos + 1], lines[pos + 2]
        pos += 3
        f = lambda line, k: float(line[15 * k: 15 * k + 15])
        high[sp, 0:5] = [f(l2, k) for k in range(5)]
        high[sp, 5] = f(l3, 0)
        high[sp, 6] = f(l3, 1)
        low[sp, 0:3] = [f(l3, k) for k in (2, 3, 4)]
        low[sp, 3:7] = [f(l4, k) for k in range(4)]
        have.add(name)
        n_read += 1
    return dict(t_low=t_low, t_mid=t_mid, t_high=t_high, low=low, high=high, mw=mw, comp=comp)


def read_therm_nasa9(therm_path, species):
    """NASA-9 reader (src/ChemData.cpp:865-1034): per species a line holding the
    range count (columns 0-1) and the molecular weight (columns 52-64), then per
    range a bounds line (columns 0-10, 11-21) and coefficient lines of 16-column
    fields a1..a5 / a6 a7 _ b1 b2."""
    ns = len(species)
    idx = {s: i for i, s in enumerate(species)}
    lines = [ln.replace("D", "E").replace("d", "e") for ln in open(therm_path).read().splitlines()]
    pos = 0
    while pos < len(lines) and "THERMO" not in lines[pos].upper():
        pos += 1
    pos += 1
    while pos < len(lines):          # default bounds header: read, unused
        s = lines[pos]
        pos += 1
        if s and s[0] != "!":
            break
    mw = np.zeros(ns)
    n_range = np.zeros(ns, dtype=int)
    cuts = np.zeros((ns, 5))
    poly = np.zeros((ns, 4, 9))
    have = set()
    while pos < len(lines):
        s = lines[pos].upper()
        pos += 1
        if not s or s[0] == "!":
            continue
        key = s.split()[0]
        if key == "END":
            continue
        if key in ("REACTIONS", "REAC", "TRANSPORT", "TRAN", "TRANS"):
            break
        if len(have) >= ns:
            continue
        name = s[:16].split()[0]
        n_part = int(lines[pos][:2])
        if name not in idx or name in have:
            pos += 1 + 3 * n_part
            continue
        sp = idx[name]
        mw[sp] = float(lines[pos][52:65])
        n_range[sp] = n_part
        for p in range(n_part):
            lb, l1, l2 = lines[pos + 1 + 3 * p], lines[pos + 2 + 3 * p], lines[pos + 3 + 3 * p]
            cuts[sp, p] = float(lb[0:11])
            cuts[sp, p + 1] = float(lb[11:22])
            poly[sp, p, 0:5] = [float(l1[16 * k:16 * k + 16]) for k in range(5)]
            poly[sp, p, 5] = float(l2[0:16])
            poly[sp, p, 6] = float(l2[16:32])
            poly[sp, p, 7] = float(l2[48:64])
            poly[sp, p, 8] = float(l2[64:80])
        pos += 1 + 3 * n_part
        have.add(name)
    assert len(have) == ns, "NASA-9 data missing for some species"
    return dict(mw=mw, n_range=n_range, t_cut=cuts, poly9=poly)


def read_two_temperature(path, species):
    """Two-temperature metadata (src/ChemData.cpp:130-232)."""
    ns = len(species)
    idx = {s: i for i, s in enumerate(species)}
    theta_v = np.zeros(ns)
    park = np.full(ns, 3.0e-21)
    lt_a = np.zeros((ns, ns))
    lt_b = np.zeros((ns, ns))
    th = [[] for _ in range(ns)]
    g = [[] for _ in range(ns)]
    cur = -1
    for line in open(path):
        for c in "#!":
            line = line.split(c)[0]
        tok = line.split()
        if not tok:
            continue
        key = tok[0].upper()
        if key == "SPECIES":
            cur = idx.get(tok[1].upper(), -2)
            continue
        if key == "END":
            cur = -1
            continue
        if cur < 0:
            continue
        if key == "THETA_V":
            theta_v[cur] = float(tok[1])
        elif key in ("PARK_SIGMA", "SIGMA", "OMEGA"):
            park[cur] = float(tok[1])
        elif key == "ELEC_THETA":
            th[cur] += [float(v) for v in tok[1:]]
        elif key == "ELEC_G":
            g[cur] += [float(v) for v in tok[1:]]
        elif key == "LT":
            j = idx.get(tok[1].upper())
            if j is not None:
                lt_a[cur, j], lt_b[cur, j] = float(tok[2]), float(tok[3])
    n_elec = np.zeros(ns, dtype=int)
    et = np.zeros((ns, 16))
    eg = np.zeros((ns, 16))
    for i in range(ns):
        if not g[i] and th[i]:
            g[i] = [1.0] * len(th[i])
        assert len(th[i]) == len(g[i])
        n_elec[i] = len(th[i])
        et[i, :n_elec[i]] = th[i]
        eg[i, :n_elec[i]] = g[i]
    return dict(two_temperature=1, theta_v=theta_v, park_sigma=park, lt_a=lt_a, lt_b=lt_b, n_elec=n_elec,
                elec_theta=et, elec_g=eg)


def _reduced_dipole(i, mu, eps_kb, sigma):
    if mu[i] == 0:
        return 0.0
    debye_to_cm_joule = 1e-18 * math.sqrt(1e-7)
    cm_to_angstrom = math.pow(1e8, 1.5)
    convert = debye_to_cm_joule * cm_to_angstrom / math.sqrt(BOLTZMANN)
    return convert * mu[i] / math.sqrt(eps_kb[i] * sigma[i] * sigma[i] * sigma[i])


def read_transport(tran_path, species, mw):
    """Transport reader and derived pair tables (src/ChemData.cpp:1037-1162)."""
    ns = len(species)
    idx = {s: i for i, s in enumerate(species)}
    geometry = np.zeros(ns, dtype=int)
    eps_kb = np.zeros(ns)
    sigma = np.zeros(ns)
    mu = np.zeros(ns)
    alpha = np.zeros(ns)
    zrot = np.zeros(ns)
    have = set()
    for raw in open(tran_path):
        s = raw.upper()
        if not s.strip() or s[0] == "!":
            continue
        key = s.split()[0]
        if key.startswith("END") or key.startswith("REAC"):
            break
        if key not in idx:
            continue
        if len(have) >= ns:
            break
        if key in have:
            continue
        t = s.split()
        sp = idx[key]
        geometry[sp] = int(t[1])
        eps_kb[sp], sigma[sp], mu[sp], alpha[sp], zrot[sp] = (float(v) for v in t[2:7])
        have.add(key)
    lj_inv = np.zeros(ns)
    vis_coeff = np.zeros(ns)
    zrot298 = np.zeros(ns)
    for sp in range(ns):
        lj_inv[sp] = 1.0 / eps_kb[sp]
        vis_coeff[sp] = 2.6693e-6 * math.sqrt(mw[sp]) / (sigma[sp] * sigma[sp])
        tri = eps_kb[sp] / 298
        # reproduced as written in the reference, src/ChemData.cpp:1083-1085
        f298 = (1 + 0.5 * PI * PI * PI * math.sqrt(tri) + (2 + 0.25 * PI * PI) * tri +
                math.sqrt(PI * PI * PI) * math.sqrt(tri * tri * tri))
        zrot298[sp] = zrot[sp] * f298
    wj = np.zeros((ns, ns))
    sq = np.zeros((ns, ns))
    for i in range(ns):
        for j in range(ns):
            wj[i, j] = math.pow(mw[j] / mw[i], 0.25)
            sq[i, j] = 1.0 / math.sqrt(8 * (1 + mw[i] / mw[j]))
    bdc = np.zeros((ns, ns))
    kbe = np.zeros((ns, ns))
    for j in range(ns):
        wjj = mw[j] * 0.5
        bdc[j, j] = 0.0188324 / (math.sqrt(wjj) * sigma[j] * sigma[j])
        kbe[j, j] = 1.0 / eps_kb[j]
        for k in range(j + 1, ns):
            wjk = mw[j] * mw[k] / (mw[j] + mw[k])
            xi = 1.0
            if (mu[j] != 0) != (mu[k] != 0):
                p, n = (j, k) if mu[j] != 0 else (k, j)
                alpha_n_star = alpha[n] / math.pow(sigma[n], 3)
                mu_p_star = _reduced_dipole(p, mu, eps_kb, sigma)
                xi = 1.0 + 0.25 * alpha_n_star * mu_p_star * mu_p_star * math.sqrt(eps_kb[p] / eps_kb[n])
            sjk = 0.5 * (sigma[j] + sigma[k]) * math.pow(xi, -1.0 / 6.0)
            bdc[j, k] = bdc[k, j] = 0.0188324 / (math.sqrt(wjk) * sjk * sjk)
            e = xi * xi * math.sqrt(eps_kb[j] * eps_kb[k])
            kbe[j, k] = kbe[k, j] = 1.0 / e
    return dict(geometry=geometry, LJ_potent_inv=lj_inv, vis_coeff=vis_coeff, ZRotF298=zrot298,
                WjDivWi_to_One4th=wj, sqrt_WiDivWjPl1Mul8=sq, binary_diffusivity_coeff=bdc,
                kb_over_eps_jk=kbe)


def read_reactions(mech_path, species):
    """Reaction reader (src/ChemData.cpp:1165-1493) for the Chemkin subset the
    shipped mechanisms use: A b Ea per line, '+M', '(+M)' with LOW/TROE,
    third-body efficiencies, DUPLICATE pairs merged, REV, '=>'."""
    ns = len(species)
    idx = {s: i for i, s in enumerate(species)}
    lines = open(mech_path).read().splitlines()
    pos = 0
    while pos < len(lines):
        t = lines[pos].split()
        pos += 1
        if t and t[0] in ("REACTIONS", "REAC"):
            break
    R = []

    def new():
        return dict(label=1, rev=0, sf=[0] * ns, sb=[0] * ns, order=0, A=0.0, b=0.0, Ea=0.0, A2=0.0,
                    b2=0.0, Ea2=0.0, tb=[1.0] * ns, ta=0.0, t1=0.0, t2=0.0, t3=0.0)

    def split_species(s, row, sign, rec):
        # '+' separated terms with an optional single-digit stoichiometric prefix
        terms = []
        cur = ""
        i = 0
        while i < len(s):
            ch = s[i]
            if ch == "+" and cur and not (i + 1 < len(s) and s[i + 1] == "+"):
                terms.append(cur)
                cur = ""
            else:
                cur += ch
            i += 1
        if cur:
            terms.append(cur)
        for term in terms:
            st = 1
            if term[0].isdigit():
                st = int(term[0])
                term = term[1:]
            row[idx[term]] += st
            rec["order"] += sign * st

    dup_open = []
    cur = None
    while pos < len(lines):
        raw = lines[pos]
        pos += 1
        s = raw.upper()
        if s.strip() == "END":
            break
        if not s.strip() or s[0] == "!":
            continue
        body = s.split("!")[0]
        if "=" in body:
            toks = body.split()
            eq = "".join(toks[:-3])
            rec = new()
            rec["A"], rec["b"], rec["Ea"] = float(toks[-3]), float(toks[-2]), float(toks[-1]) / R_C
            if "+M" in eq:
                rec["label"] = 4
                if "(+" in eq:
                    rec["label"] = 5
                    eq = eq.replace("(+M)", "")
                else:
                    eq = eq.replace("+M", "", 2)
            if "<=>" in eq:
                lhs, rhs = eq.split("<=>")
            elif "=>" in eq:
                lhs, rhs = eq.split("=>")
                rec["label"] = 0
            else:
                lhs, rhs = eq.split("=")
            split_species(lhs, rec["sf"], -1, rec)
            split_species(rhs, rec["sb"], +1, rec)
            R.append(rec)
            cur = rec
            continue
        # auxiliary line for the current reaction
        aux = body.replace("/", " ").split()
        if not aux:
            continue
        key = aux[0]
        if key in ("DUPLICATE", "DUP"):
            cur["label"] = 3
            cur["dup"] = True
        elif key == "REV":
            cur["A2"], cur["b2"], cur["Ea2"] = float(aux[1]), float(aux[2]), float(aux[3]) / R_C
            cur["rev"] = 1
        elif key == "LOW":
            cur["A2"], cur["b2"], cur["Ea2"] = float(aux[1]), float(aux[2]), float(aux[3]) / R_C
        elif key == "TROE":
            cur["label"] = 6
            cur["ta"], cur["t3"], cur["t1"] = float(aux[1]), float(aux[2]), float(aux[3])
            if len(aux) > 4:
                cur["t2"] = float(aux[4])
                cur["label"] = 7
        else:
            for a in range(0, len(aux) - 1, 2):
                cur["tb"][idx[aux[a]]] = float(aux[a + 1])
    # merge DUPLICATE pairs: the second's Arrhenius set goes to A2/b2/Ea2 of the first
    out = []
    pending = []
    for rec in R:
        if rec.get("dup"):
            match = None
            for k, prev in enumerate(pending):
                if prev["sf"] == rec["sf"] and prev["sb"] == rec["sb"]:
                    match = k
                    break
            if match is not None:
                prev = pending.pop(match)
                prev["A2"], prev["b2"], prev["Ea2"] = rec["A"], rec["b"], rec["Ea"]
                continue
            pending.append(rec)
        out.append(rec)
    return out


def build_tables(mech_path, therm_path, tran_path, nasa=7, two_temperature=None):
    """All tables as a dict of numpy arrays (species order = mechanism order)."""
    elems, species = read_species(mech_path)
    ns = len(species)
    if nasa == 9:
        th9 = read_therm_nasa9(therm_path, species)
        mw = th9["mw"]
        th = dict(t_low=np.zeros(ns), t_mid=np.zeros(ns), t_high=np.zeros(ns), low=np.zeros((ns, 7)),
                  high=np.zeros((ns, 7)))
    else:
        th = read_therm_nasa7(therm_path, species, elems)
        mw = th["mw"]
        th9 = dict(n_range=np.zeros(ns, dtype=int), t_cut=np.zeros((ns, 5)), poly9=np.zeros((ns, 4, 9)))
    tr = read_transport(tran_path, species, mw)
    rx = read_reactions(mech_path, species)
    nr = len(rx)
    T = dict(species=species, n_spec=ns, n_reac=nr, nasa7or9=nasa, mw=mw, imw=1.0 / mw, gas_const=R_U / mw,
             t_low=th["t_low"], t_mid=th["t_mid"], t_high=th["t_high"], low_coeff=th["low"],
             high_coeff=th["high"], n_range=th9["n_range"], t_cut=th9["t_cut"], poly9=th9["poly9"])
    T.update(tr)
    T["reac_type"] = np.array([r["label"] for r in rx], dtype=int)
    T["rev_type"] = np.array([r["rev"] for r in rx], dtype=int)
    T["order"] = np.array([r["order"] for r in rx], dtype=int)
    T["stoi_f"] = np.array([r["sf"] for r in rx], dtype=int)
    T["stoi_b"] = np.array([r["sb"] for r in rx], dtype=int)
    for k in ("A", "b", "Ea", "A2", "b2", "Ea2"):
        T[k] = np.array([r[k] for r in rx])
    T["third_body"] = np.array([r["tb"] for r in rx])
    T["troe_alpha"] = np.array([r["ta"] for r in rx])
    T["troe_t1"] = np.array([r["t1"] for r in rx])
    T["troe_t2"] = np.array([r["t2"] for r in rx])
    T["troe_t3"] = np.array([r["t3"] for r in rx])
    if two_temperature:
        T.update(read_two_temperature(two_temperature, species))
    else:
        T.update(two_temperature=0, theta_v=np.zeros(ns), park_sigma=np.zeros(ns), lt_a=np.zeros((ns, ns)),
                 lt_b=np.zeros((ns, ns)), n_elec=np.zeros(ns, dtype=int), elec_theta=np.zeros((ns, 16)),
                 elec_g=np.zeros((ns, 16)))
    # reaction temperatures T^a Tve^b: a = 1, b = 0 by default (src/ChemData.cpp:1294-1297); the
    # two-temperature N2/O2/NO/N/O 17-reaction mechanism uses sqrt(T Tve) for its first 15
    T["trx_model"] = 0
    T["tcf_a"], T["tcb_a"] = np.ones(nr), np.ones(nr)
    T["tcf_b"], T["tcb_b"] = np.zeros(nr), np.zeros(nr)
    if two_temperature and ns == 5 and nr == 17 and set(species) == {"N2", "O2", "NO", "N", "O"}:
        T["trx_model"] = 1
        T["tcf_a"][:15] = 0.5
        T["tcf_b"][:15] = 0.5
    return T
