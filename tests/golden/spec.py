"""What the golden fixtures hold: limiters, ghost modes and the assembled operators (the
terms of dare_terms / orc_terms as keyword arguments). Shared by make_golden.py and the
tests that read the fixtures."""


LIMITERS = {"cds": 0, "upwind": 1, "vanalbada": 2, "minmod": 3, "vanleer": 4}
GHOST_MODES = [2, 3, 4, 2, 3, 4]
GHOST_VALUES = [0.5, 0.0, -1.25, 2.0, 0.0, 0.75]


def operators(n):
    """(name, ncomp, terms keywords, extra arrays) of the assembled operators at size n."""
    ops = [("c2diff", 1, dict(dt=1e-2, diff=1e-3, bc_type=[1] * 6, bc_value=[0.0] * 6), ())]
    if n == 8:
        ops.append(("varconv", 1, dict(dt=5e-3, limiter=4, bc_type=[1, 2, 1, 2, 1, 1],
                                       bc_value=[0.3, 0.0, -1.0, 0.0, 0.5, 0.25]), ("diff_cell", "source", "vel")))
        ops.append(("react2", 2, dict(dt=0.1, diff_comp=[1e-3, 5e-3], bc_type=[1] * 6,
                                      bc_value=[0.1, -0.2, 0.3, 0.0, 0.0, 0.5]), ("react",)))
    return ops
