"""Write golden coefficient vectors for compi from the reference's own formulas.

The reference's coefficient kernel (compi.cu, coef3D and coef2D) cannot be
built here (it needs nvcc and an NVIDIA GPU).  Its per-pair arithmetic is a
straight-line list of C assignments, so this script reads compi.cu as text
and evaluates those statements for a small random problem:

  * the pair-weight block (compi.cu:747-791 for 3-D, 193-228 for 2-D) runs
    statement by statement with C semantics for the declared types: double
    variables keep FP64, assignments to float variables round to FP32, and
    ``detX[..]*detY[..]`` is an FP32 product;
  * every ``temp=...`` expression (compi.cu:800-3350 / 236-417) is evaluated
    exactly in FP64 on those FP32 values and accumulated as ``P*temp`` into
    slot ``i``, which is what the shared-memory atomics sum.

The staging of compi.cu:642-677 (X_C*Isig, det*sqrtf(Isig^3) in 3-D;
compi.cu:106-148 with the round-down Isig in 2-D) is applied with numpy FP32.

Output: compi_golden.npz beside this script (inputs and expected vectors).
Run where the reference is present:  python tests/golden/make_compi_golden.py
"""
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def kernel_text(src, name, end_marker):
    start = src.index("__global__ void %s(float *d_X" % name)
    start = src.index("{", start)
    end = src.index(end_marker, start)
    return src[start:end]


def split_kernel(body):
    """Return (pair statements, [(slot, temp expression)])."""
    a = body.index("if (ix<N && iy<M)")
    a = body.index("{", a) + 1
    first_slot = body.index("i=0;", a)
    pair_src = body[a:first_slot]
    slots = re.findall(r"\bi=(\d+);\s*temp=\s*([^;]*);", body[first_slot:])
    return pair_src, [(int(i), e) for i, e in slots]


def f32(x):
    return float(np.float32(x))


def run_pair(pair_src, env):
    """Execute the pair-weight statements with C type semantics."""
    floats = set()
    code = re.sub(r"//[^\n]*", "", pair_src)
    for stmt in code.split(";"):
        s = " ".join(stmt.split())
        if not s or s.startswith("__syncthreads") or re.match(r"^int\b", s) \
                or s.startswith("double temp"):
            continue
        m = re.match(r"^(float|double)\s+(.*)$", s)
        if m:
            kind, rest = m.groups()
            for decl in rest.split(","):
                decl = decl.strip()
                if "=" in decl:
                    name, val = decl.split("=", 1)
                    env[name.strip()] = float(eval(val, {}, {}))
                    name = name.strip()
                else:
                    name = decl
                if kind == "float":
                    floats.add(name)
            continue
        m = re.match(r"^(\w+)\s*=\s*(.*)$", s)
        if not m:
            raise ValueError("unhandled statement: %r" % s)
        name, expr = m.groups()
        if re.search(r"\b(X|Y|X_C|Y_C|d_Tran_Coef)\[", expr):
            # loads from staged shared memory / constants: provided by env
            continue
        expr = expr.replace("detX[threadIdx.x]*detY[threadIdx.y]", "f32(detX*detY)")
        expr = expr.replace("exp(", "math_exp(")
        val = eval(expr, {"f32": f32, "math_exp": np.exp}, env)
        env[name] = f32(val) if name in floats else float(val)
    return env


def staged_3d(inv, g, isig):
    """compi.cu:655-677: X_C*Isig in FP32, det*sqrtf(Isig^3) in FP32."""
    isig = np.float32(isig)
    s = np.sqrt(np.float32(isig * isig * isig), dtype=np.float32)
    return (inv.astype(np.float32) * isig).astype(np.float32), (g.astype(np.float32) * s).astype(np.float32)


def staged_2d(inv, g, isig):
    """compi.cu:119-147: X_C*Isig and det*Isig in FP32."""
    isig = np.float32(isig)
    return (inv.astype(np.float32) * isig).astype(np.float32), (g.astype(np.float32) * isig).astype(np.float32)


def isig_value(D, sig):
    inv = 1.0 / sig
    f = np.float32(inv)
    if D == 2 and float(f) > inv:  # __double2float_rd (compi.cu:110)
        f = np.nextafter(f, np.float32(-np.inf))
    return f


def names_for(D):
    if D == 3:
        # column-major 3x3: entry k -> a_{(k%3)+1, (k//3)+1}
        mat = ["%d%d" % (k % 3 + 1, k // 3 + 1) for k in range(9)]
    else:
        mat = ["%d%d" % (k % 2 + 1, k // 2 + 1) for k in range(4)]
    return mat


def golden(D, N, M, seed, sig, src):
    rng = np.random.default_rng(seed)
    if D == 3:
        body = kernel_text(src, "coef3D", "__syncthreads();     \n    \n     //check")
    else:
        body = kernel_text(src, "coef2D", "__global__ void compi_redu")
    pair_src, slots = split_kernel(body)
    nslots = 637 if D == 3 else 45
    assert sorted(i for i, _ in slots) == list(range(nslots))

    Xf = (rng.normal(size=(N, D)) * 0.4).astype(np.float32)
    Yf = (rng.normal(size=(M, D)) * 0.4).astype(np.float32)

    def spd(n):
        A = rng.normal(size=(n, D, D)) * 0.5
        return A @ np.transpose(A, (0, 2, 1)) + 0.3 * np.eye(D)

    def inverse_state(C):
        inv = np.linalg.inv(C.astype(np.float32).astype(np.float64)).astype(np.float32)
        inv = np.transpose(inv, (0, 2, 1)).reshape(len(C), D * D)  # column-major
        g = np.sqrt(1.0 / np.linalg.det(C)).astype(np.float32)
        return np.ascontiguousarray(inv), g

    XCinv, gX = inverse_state(spd(N))
    YCinv, gY = inverse_state(spd(M))
    ang = rng.uniform(-0.5, 0.5, size=3)
    if D == 3:
        K = np.array([[0, -ang[2], ang[1]], [ang[2], 0, -ang[0]], [-ang[1], ang[0], 0]])
        th = np.linalg.norm(ang)
        R = np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K
    else:
        c, s = np.cos(ang[0]), np.sin(ang[0])
        R = np.array([[c, -s], [s, c]])
    t = rng.normal(size=D) * 0.2

    isig = isig_value(D, sig)
    stage = staged_3d if D == 3 else staged_2d
    Xs, detXs = stage(XCinv, gX, isig)
    Ys, detYs = stage(YCinv, gY, isig)
    mat = names_for(D)
    Rc = R.T.ravel()  # column-major, as d_Tran_Coef holds it
    coef = np.zeros(nslots)
    for i in range(N):
        for j in range(M):
            env = {}
            for a in range(D):
                env["n%d" % (a + 1)] = float(Xf[i, a])
                env["m%d" % (a + 1)] = float(Yf[j, a])
                env["t%d" % (a + 1)] = float(t[a])
            for k, nm in enumerate(mat):
                env["a" + nm] = float(Ys[j, k])
                env["b" + nm] = float(Xs[i, k])
                env["r" + nm] = float(Rc[k])
            env["detX"] = float(detXs[i])
            env["detY"] = float(detYs[j])
            run_pair(pair_src, env)
            P = env["P"]
            for slot, expr in slots:
                coef[slot] += P * float(eval(expr, {}, env))
    return dict(Xf=Xf, XCinv=XCinv, gX=gX, Yf=Yf, YCinv=YCinv, gY=gY, R=R, t=t,
                sig=np.float64(sig), coef=coef)


def main():
    src = open(os.path.join(REF, "compi.cu")).read()
    g3 = golden(3, 9, 7, 11, 0.8, src)
    g2 = golden(2, 8, 6, 12, 0.6, src)
    out = {}
    for k, v in g3.items():
        out["d3_" + k] = v
    for k, v in g2.items():
        out["d2_" + k] = v
    dst = os.path.join(HERE, "compi_golden.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst)


if __name__ == "__main__":
    main()
