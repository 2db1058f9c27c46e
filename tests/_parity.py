"""Helpers shared by the GPU parity tests: bit-for-bit comparison of the
libhydist state against the CPU checker."""
import numpy as np

STATE_F64 = ("u", "v", "z", "t_u", "t_v", "t_z", "H_moi", "Htdu", "Htdv", "htaiz", "htaiz_bd", "Kx1", "Ky1",
             "FS", "tFS", "VTH", "Kx", "Ky", "Qbx", "Qby", "dH", "hi", "ubt", "ubp", "vbt", "vbd")
STATE_INT = ("khouot", "moci", "mocj")


def diff_fields(a_get, b_get, names=STATE_F64 + STATE_INT):
    """Names whose arrays differ (NaN equal to NaN), with a short report."""
    out = []
    for n in names:
        a, b = a_get(n), b_get(n)
        eq = (a == b)
        if a.dtype.kind == "f":
            eq |= np.isnan(a) & np.isnan(b)
        if not eq.all():
            idx = np.argwhere(~eq)
            out.append(f"{n}: {int((~eq).sum())} differ, first at {idx[0].tolist()}: "
                       f"{a[tuple(idx[0])]!r} vs {b[tuple(idx[0])]!r}")
    return out


def diff_segments(a_get, b_get):
    """Segment tables compared on their valid entries (first moc per line)."""
    out = []
    for mo, da, cu in (("moci", "daui", "cuoii"), ("mocj", "dauj", "cuoij")):
        m1, m2 = a_get(mo), b_get(mo)
        d1, d2, c1, c2 = a_get(da), b_get(da), a_get(cu), b_get(cu)
        bad = [l for l in range(m1.size)
               if m1[l] != m2[l] or (d1[l, :m1[l]] != d2[l, :m1[l]]).any() or (c1[l, :m1[l]] != c2[l, :m1[l]]).any()]
        if bad:
            out.append(f"{da}/{cu}: {len(bad)} lines differ, first {bad[0]}")
    return out


def assert_same_state(o, g, where=""):
    d = diff_fields(o.get, g.get) + diff_segments(o.get, g.get)
    assert not d, f"{where}: " + "; ".join(d[:6])
