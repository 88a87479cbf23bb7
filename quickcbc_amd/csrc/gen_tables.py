"""Extract the numerical tables that the waveform path must share with QuickCBC.

Three tables are data, not algorithm, and must be identical to QuickCBC's:

* the Kerr quasi-normal-mode table (1003 spins, ringdown and damping
  frequencies) that IMRPhenomD splines for f_RD and f_damp
  (IMRPhenomD_internals.h:176-617);
* the antenna-pattern polynomials F+ / Fx of the three detectors, written in
  QuickCBC as expanded trigonometric polynomials with 10-digit coefficients
  (Utilities.c:1004-1076);
* the detector vertex vectors used for geocentre-to-detector light-travel
  delays (Utilities.c:1078-1109).

This script reads those source files as text and writes ``qcbc_tables.inc``
(hex-float literals, so the committed values are exactly the doubles a C
compiler would produce from the decimal literals).  It is run once in a
checkout that has the QuickCBC sources; its output is committed.

Usage: python gen_tables.py /path/to/QuickCBC [out.inc]
"""
import os
import re
import sys

NUM = r"(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?"
TERM = re.compile(r"([+-]?)\s*(" + NUM + r")((?:\s*\*\s*[A-Za-z_][A-Za-z0-9_]*)*)")

# factor names used by the antenna polynomials, and how each maps to the
# basic variables (index, power)
ANT_VARS = ["cpsi", "spsi", "cdelta", "sdelta", "calpha", "salpha", "cpsi2", "sdelta2", "calpha2"]
ANT_ALIAS = {}
TIME_VARS = ["cosd", "calpha", "salpha", "sind"]


def parse_array(text, name):
    m = re.search(r"static const double " + name + r"\[\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise SystemExit("array %s not found" % name)
    body = m.group(1).replace("\\", " ")
    vals = [float(t) for t in re.findall(r"[+-]?" + NUM, body)]
    return vals


def parse_poly(expr, varnames, alias):
    """Split an expanded polynomial into (coef, [factor indices in source order])."""
    expr = expr.replace("\n", " ").strip()
    pos = 0
    terms = []
    while pos < len(expr):
        m = TERM.match(expr, pos)
        if not m or m.end() == pos:
            raise SystemExit("cannot parse polynomial near: %r" % expr[pos:pos + 40])
        sign, num, facs = m.groups()
        coef = float(num) * (-1.0 if sign == "-" else 1.0)
        idx = []
        for f in re.findall(r"[A-Za-z_][A-Za-z0-9_]*", facs):
            if f in alias:
                base, pw = alias[f]
                idx.extend([varnames.index(base)] * pw)
            else:
                idx.append(varnames.index(f))
        terms.append((coef, idx, facs))
        pos = m.end()
        while pos < len(expr) and expr[pos] == " ":
            pos += 1
    return terms


FITS = ["rho1_fun", "rho2_fun", "rho3_fun", "gamma1_fun", "gamma2_fun", "gamma3_fun",
        "AmpIntColFitCoeff", "alpha1Fit", "alpha2Fit", "alpha3Fit", "alpha4Fit", "alpha5Fit",
        "beta1Fit", "beta2Fit", "beta3Fit", "sigma1Fit", "sigma2Fit", "sigma3Fit", "sigma4Fit"]


def fit_numbers(src, name, pattern):
    """Signed numeric literals of a fit's return expression, in source order.

    ``pattern`` is the expected sequence of variable names in the expression,
    which guards the coefficient layout the C code assumes."""
    i = src.index("double " + name + "(")
    j = src.index("return", i)
    expr = src[j + 6:src.index(";", j)]
    names = re.findall(r"\b(eta[234]?|xi[23]?|s[234]?)\b", expr)
    if names != pattern:
        raise SystemExit("unexpected layout in %s: %s" % (name, names))
    expr = re.sub(r"\b(eta[234]?|xi[23]?|s[234]?)\b", "V", expr)
    toks = re.findall(r"([+-]?)\s*(" + NUM + r")", expr)
    return [float(n) * (-1.0 if sg == "-" else 1.0) for sg, n in toks]


def grab(text, func, lhs):
    start = text.index("void " + func + "(")
    m = re.compile(re.escape(lhs) + r"\s*=\s*(.*?);", re.S).search(text, start)
    return m.group(1)


def main():
    ref = sys.argv[1]
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(os.path.dirname(os.path.abspath(__file__)), "qcbc_tables.inc")
    hdr = open(os.path.join(ref, "IMRPhenomD_internals.h")).read()
    util = open(os.path.join(ref, "Utilities.c")).read()
    src = open(os.path.join(ref, "IMRPhenomD_internals.c")).read()

    qa = parse_array(hdr, "QNMData_a")
    qr = parse_array(hdr, "QNMData_fring")
    qd = parse_array(hdr, "QNMData_fdamp")
    assert len(qa) == len(qr) == len(qd) == 1003, (len(qa), len(qr), len(qd))
    assert all(qa[i] < qa[i + 1] for i in range(len(qa) - 1))

    polys = []
    for det, func in enumerate(["F_LHO", "F_LLO", "F_VIRGO"]):
        for comp, lhs in enumerate(["*Fplus", "*Fcross"]):
            terms = parse_poly(grab(util, func, lhs), ANT_VARS, ANT_ALIAS)
            polys.append((det, comp, terms))

    tstart = util.index("void Times(")
    tbody = util[tstart:util.index("}", tstart)]
    verts = []
    for name in ["LHO", "LLO", "VIRGO"]:
        m = re.search(r"\n\s*" + name + r"\s*=\s*(.*?);", tbody, re.S)
        terms = parse_poly(m.group(1), TIME_VARS, {})
        verts.append(terms)

    L = []
    L.append("/* Generated by gen_tables.py from the QuickCBC sources; do not edit.")
    L.append("   QCBC_TABLE_STORAGE: tables indexed at run time (the QNM spline);")
    L.append("   QCBC_CONST_STORAGE: tables whose indices are compile-time constants. */")
    L.append("#ifndef QCBC_CONST_STORAGE")
    L.append("#define QCBC_CONST_STORAGE QCBC_TABLE_STORAGE")
    L.append("#endif")
    L.append("/* Kerr QNM table, IMRPhenomD_internals.h:176-617 (1003 spins). */")
    L.append("#define QCBC_QNM_N 1003")
    for nm, arr in (("qcbc_qnm_a", qa), ("qcbc_qnm_fring", qr), ("qcbc_qnm_fdamp", qd)):
        L.append("QCBC_TABLE_STORAGE double %s[QCBC_QNM_N] = {" % nm)
        for i in range(0, len(arr), 4):
            L.append("  " + ", ".join(float.hex(v) for v in arr[i:i + 4]) + ",")
        L.append("};")
    L.append("/* Antenna-pattern polynomials, Utilities.c:1004-1076.  Term t of")
    L.append("   polynomial p is coef * x[f0] * x[f1] ... (factors applied left to right")
    L.append("   in source order, -1 ends the list); x = {cos psi, sin psi, cos delta,")
    L.append("   sin delta, cos(alpha-GMST), sin(alpha-GMST), cos^2 psi, sin^2 delta,")
    L.append("   cos^2(alpha-GMST)} with the squares formed first.  Polynomial index is")
    L.append("   2*detector + (0 for F+, 1 for Fx); detectors 0=H, 1=L, 2=V. */")
    nt_max = max(len(t) for _, _, t in polys)
    L.append("#define QCBC_ANT_MAXTERMS %d" % nt_max)
    L.append("#define QCBC_ANT_MAXFAC 6")
    L.append("QCBC_CONST_STORAGE int qcbc_ant_nterms[6] = {%s};" % ", ".join(str(len(t)) for _, _, t in polys))
    L.append("QCBC_CONST_STORAGE double qcbc_ant_coef[6][QCBC_ANT_MAXTERMS] = {")
    for det, comp, terms in polys:
        row = [float.hex(c) for c, _, _ in terms] + ["0x0p+0"] * (nt_max - len(terms))
        L.append("  {" + ", ".join(row) + "},")
    L.append("};")
    L.append("QCBC_CONST_STORAGE signed char qcbc_ant_fac[6][QCBC_ANT_MAXTERMS][QCBC_ANT_MAXFAC] = {")
    for det, comp, terms in polys:
        rows = []
        for _, idx, _ in terms:
            assert len(idx) <= 6
            rows.append("{" + ", ".join(str(i) for i in idx + [-1] * (6 - len(idx))) + "}")
        rows += ["{-1, -1, -1, -1, -1, -1}"] * (nt_max - len(terms))
        L.append("  {" + ", ".join(rows) + "},")
    L.append("};")
    L.append("/* Phenomenological fits of IMRPhenomD (arXiv:1508.07253 Table 5), as")
    L.append("   c0 + c1*eta + (c2 + c3*eta + c4*eta^2)*xi + (c5 + c6*eta + c7*eta^2)*xi^2")
    L.append("   + (c8 + c9*eta + c10*eta^2)*xi^3 with xi = chiPN - 1.  Source:")
    L.append("   IMRPhenomD_internals.c:552-592 (rho), :712-752 (gamma), :827-837 (v2),")
    L.append("   :1147-1217 (alpha), :1253-1293 (beta), :1335-1390 (sigma). */")
    fitpat = ["eta", "eta", "eta2", "xi", "eta", "eta2", "xi2", "eta", "eta2", "xi3"]
    L.append("enum { " + ", ".join("QCBC_FIT_%s" % ("V2" if f == "AmpIntColFitCoeff" else f.replace("_fun", "").replace("Fit", "").upper()) for f in FITS) + ", QCBC_NFIT };")
    L.append("QCBC_CONST_STORAGE double qcbc_fit[QCBC_NFIT][11] = {")
    for f in FITS:
        v = fit_numbers(src, f, fitpat)
        assert len(v) == 11, (f, v)
        L.append("  {" + ", ".join(float.hex(x) for x in v) + "}, /* %s */" % f)
    L.append("};")
    fs = fit_numbers(src, "FinalSpin0815_s", ["eta", "eta2", "eta3", "eta4", "eta", "eta2", "s", "eta", "eta2", "s2", "eta", "eta2", "s3", "eta", "eta2", "s4"])
    assert len(fs) == 13
    L.append("/* Final spin fit, arXiv:1508.07250 Eq. 3.6; IMRPhenomD_internals.c:415-430:")
    L.append("   c0 eta + c1 eta^2 + c2 eta^3 + c3 eta^4 + (c4 + c5 eta + c6 eta^2) s")
    L.append("   + (c7 eta + c8 eta^2) s^2 + (c9 eta + c10 eta^2) s^3 + (c11 eta + c12 eta^2) s^4 */")
    L.append("QCBC_CONST_STORAGE double qcbc_finalspin[13] = {" + ", ".join(float.hex(x) for x in fs) + "};")
    er = fit_numbers(src, "EradRational0815_s", ["eta", "eta2", "eta3", "eta4", "eta", "eta2", "s", "eta", "eta2", "s"])
    assert len(er) == 12
    L.append("/* Radiated energy fit, arXiv:1508.07250 Eqs. 3.7-3.8; IMRPhenomD_internals.c:450-459:")
    L.append("   (c0 eta + c1 eta^2 + c2 eta^3 + c3 eta^4)(c4 + (c5 + c6 eta + c7 eta^2) s)")
    L.append("   / (c8 + (c9 + c10 eta + c11 eta^2) s) */")
    L.append("QCBC_CONST_STORAGE double qcbc_erad[12] = {" + ", ".join(float.hex(x) for x in er) + "};")
    L.append("/* Detector vertex projections, Utilities.c:1078-1109: r_d . n =")
    L.append("   c0*cosd*cos(GMST-alpha) + c1*cosd*sin(GMST-alpha) + c2*sind (metres). */")
    L.append("QCBC_CONST_STORAGE double qcbc_vertex[3][3] = {")
    for terms in verts:
        assert [t[1] for t in terms] == [[0, 1], [0, 2], [3]], terms
        L.append("  {" + ", ".join(float.hex(t[0]) for t in terms) + "},")
    L.append("};")
    open(out, "w").write("\n".join(L) + "\n")
    print("wrote", out, "qnm", len(qa), "antenna terms", [len(t) for _, _, t in polys])


if __name__ == "__main__":
    main()
