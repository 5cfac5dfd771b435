"""Field and source component names (same strings as emopt/defs.py:11-25)."""


class FieldComponent:
    Ex = "Ex"
    Ey = "Ey"
    Ez = "Ez"
    Hx = "Hx"
    Hy = "Hy"
    Hz = "Hz"


class SourceComponent:
    Jx = "Jx"
    Jy = "Jy"
    Jz = "Jz"
    Mx = "Mx"
    My = "My"
    Mz = "Mz"


FIELD_INDEX = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}
