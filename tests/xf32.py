"""Small float32 Transform helpers for building test fixtures (translate,
scale, product with m / m_inv as transform.rs:40-65,266-272 define them)."""
import numpy as np


def translate(v):
    m = np.eye(4, dtype=np.float32)
    mi = np.eye(4, dtype=np.float32)
    m[:3, 3] = np.asarray(v, np.float32)
    mi[:3, 3] = -np.asarray(v, np.float32)
    return m, mi


def scale(x, y, z):
    m = np.diag(np.array([x, y, z, 1], np.float32))
    mi = np.diag(np.array([np.float32(1) / np.float32(x), np.float32(1) / np.float32(y),
                           np.float32(1) / np.float32(z), 1], np.float32))
    return m, mi


def matmul(a, b):
    """Matrix4x4 product, matrix4x4.rs:167-182, summed left to right in f32."""
    r = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            s = np.float32(a[i, 0] * b[0, j])
            s = np.float32(s + np.float32(a[i, 1] * b[1, j]))
            s = np.float32(s + np.float32(a[i, 2] * b[2, j]))
            s = np.float32(s + np.float32(a[i, 3] * b[3, j]))
            r[i, j] = s
    return r


def mul(t1, t2):
    return matmul(t1[0], t2[0]), matmul(t2[1], t1[1])


def xf_vector(m, v):
    v = np.asarray(v, np.float32)
    out = []
    for i in range(3):
        s = np.float32(m[i, 0] * v[0])
        s = np.float32(s + np.float32(m[i, 1] * v[1]))
        s = np.float32(s + np.float32(m[i, 2] * v[2]))
        out.append(s)
    return np.array(out, np.float32)
