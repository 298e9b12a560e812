"""Test infrastructure: a float64 restatement of the reference interpreter.

Evaluates ProgramBuffer words (tangerine/sdf_evaluator.h:83-107) the way
SDFInterpreter::Eval does (tangerine/sdf_evaluator.cpp:1384-1605), with the
SDFMath formulas of :165-295, in float64 and vectorised over points with
numpy.  The float32 words (matrix entries, parameters, thresholds) are taken
as they are, so the result is the program's own function evaluated with
rounding errors about 2^29 times smaller than the float32 interpreter's: the
float32 interpreter's rounding error at a point is its difference from this
value.  Used only by tests (the pruning slack's error constant).

Pinned by tests/test_oracle.py: on the golden fixtures it agrees with the
reference's float32 interpreter values to within float32 rounding, and on
closed forms (sphere |p| - r, box faces, blends at a = b) to float64
rounding.
"""

from __future__ import annotations

import numpy as np

STOP, SPHERE, ELLIPSOID, BOX, TORUS, CYLINDER, CONE, CONINDER, PLANE = range(9)
UNION, INTER, DIFF, BLEND_UNION, BLEND_INTER, BLEND_DIFF, FLATE, OFFSET, MATRIX, SCALE_FIELD = range(9, 19)
_NPARAMS = {SPHERE: 1, ELLIPSOID: 3, BOX: 3, TORUS: 2, CYLINDER: 2, CONE: 2, CONINDER: 3, PLANE: 3}


def _len2(x, y):
    return np.sqrt(x * x + y * y)


def _brush(op, prm, x, y, z):
    """SDFMath primitives (sdf_evaluator.cpp:165-248) in float64."""
    if op == SPHERE:
        return np.sqrt(x * x + y * y + z * z) - prm[0]
    if op == ELLIPSOID:
        k0 = np.sqrt((x / prm[0]) ** 2 + (y / prm[1]) ** 2 + (z / prm[2]) ** 2)
        k1 = np.sqrt((x / prm[0] ** 2) ** 2 + (y / prm[1] ** 2) ** 2 + (z / prm[2] ** 2) ** 2)
        return k0 * (k0 - 1.0) / k1
    if op == BOX:
        ax, ay, az = np.abs(x) - prm[0], np.abs(y) - prm[1], np.abs(z) - prm[2]
        m = np.sqrt(np.maximum(ax, 0) ** 2 + np.maximum(ay, 0) ** 2 + np.maximum(az, 0) ** 2)
        return m + np.minimum(np.maximum(np.maximum(ax, ay), az), 0.0)
    if op == TORUS:
        return _len2(_len2(x, y) - prm[0], z) - prm[1]
    if op == CYLINDER:
        dx, dz = np.abs(_len2(x, y)) - prm[0], np.abs(z) - prm[1]
        return np.minimum(np.maximum(dx, dz), 0.0) + _len2(np.maximum(dx, 0.0), np.maximum(dz, 0.0))
    if op == PLANE:
        return x * prm[0] + y * prm[1] + z * prm[2]
    if op == CONE:
        tangent, height = prm[0], prm[1]
        qx, qy = height * tangent, -height
        wx, wy = _len2(x, y), height * -0.5 + z
        t0 = np.clip((wx * qx + wy * qy) / (qx * qx + qy * qy), 0.0, 1.0)
        ax, ay = wx - qx * t0, wy - qy * t0
        t1 = np.clip(wx / qx, 0.0, 1.0)
        bx, by = wx - qx * t1, wy - qy
        k = np.sign(qy)
        d = np.minimum(ax * ax + ay * ay, bx * bx + by * by)
        s = np.maximum(k * (wx * qy - wy * qx), k * (wy - qy))
        return np.sqrt(d) * np.sign(s)
    if op == CONINDER:
        rl, rh, h = prm
        qx, qy = _len2(x, y), z
        k2x, k2y = rh - rl, 2.0 * h
        cax = qx - np.minimum(qx, np.where(qy < 0.0, rl, rh))
        cay = np.abs(qy) - h
        t = np.clip(((rh - qx) * k2x + (h - qy) * k2y) / (k2x * k2x + k2y * k2y), 0.0, 1.0)
        cbx, cby = (qx - rh) + k2x * t, (qy - h) + k2y * t
        s = np.where((cbx < 0.0) & (cay < 0.0), -1.0, 1.0)
        return s * np.sqrt(np.minimum(cax * cax + cay * cay, cbx * cbx + cby * cby))
    raise ValueError(f"not a brush: {op}")


def _set(op, a, b, t):
    """Set operators (sdf_evaluator.cpp:252-293) in float64."""
    if op == UNION:
        return np.minimum(a, b)
    if op == INTER:
        return np.maximum(a, b)
    if op == DIFF:
        return np.maximum(a, -b)
    if op == BLEND_UNION:
        h = np.maximum(t - np.abs(a - b), 0.0)
        return np.minimum(a, b) - h * h * 0.25 / t
    if op == BLEND_INTER:
        h = np.maximum(t - np.abs(a - b), 0.0)
        return np.maximum(a, b) + h * h * 0.25 / t
    if op == BLEND_DIFF:
        h = np.maximum(t - np.abs(a + b), 0.0)
        return np.maximum(a, -b) + h * h * 0.25 / t
    raise ValueError(f"not a set operator: {op}")


def interp64(words, points) -> np.ndarray:
    """The program's value at each point, in float64."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    f = w.view(np.float32).astype(np.float64)
    p = np.asarray(points, dtype=np.float32).reshape(-1, 3).astype(np.float64)
    ex, ey, ez = p[:, 0], p[:, 1], p[:, 2]
    x, y, z = ex, ey, ez
    stack: list[np.ndarray] = []
    pc = 0
    while pc < w.size:
        op = int(w[pc])
        pc += 1
        if op == STOP:
            if len(stack) != 1:
                raise ValueError("stack not balanced at Stop")
            return stack[0]
        if op in _NPARAMS:
            n = _NPARAMS[op]
            stack.append(_brush(op, f[pc:pc + n], x, y, z))
            pc += n
            x, y, z = ex, ey, ez
        elif UNION <= op <= BLEND_DIFF:
            t = 0.0
            if op >= BLEND_UNION:
                t = f[pc]
                pc += 1
            b = stack.pop()
            a = stack.pop()
            stack.append(_set(op, a, b, t))
        elif op == FLATE:
            stack[-1] = stack[-1] - f[pc]
            pc += 1
        elif op == SCALE_FIELD:
            stack[-1] = stack[-1] * f[pc]
            pc += 1
        elif op == OFFSET:
            x, y, z = ex + f[pc], ey + f[pc + 1], ez + f[pc + 2]
            pc += 3
        elif op == MATRIX:
            m = f[pc:pc + 16]  # column-major
            pc += 16
            x = m[0] * ex + m[4] * ey + m[8] * ez + m[12]
            y = m[1] * ex + m[5] * ey + m[9] * ez + m[13]
            z = m[2] * ex + m[6] * ey + m[10] * ez + m[14]
        else:
            raise ValueError(f"unknown opcode {op} at word {pc - 1}")
    raise ValueError("program without Stop")
