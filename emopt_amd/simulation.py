"""Abstract solver surface, the contract of emopt/simulation.py:20-175.

:class:`emopt_amd.fdtd.FDTD` implements every abstract method, so code written
against EMopt's ``MaxwellSolver`` (adjoint methods, optimisers) can hold an
instance of it unchanged.
"""
from abc import ABC, abstractmethod


class MaxwellSolver(ABC):
    def __init__(self, ndims):
        self._field_domains = []
        self._saved_fields = []
        self._source_power = 0.0
        self._ndims = ndims

    @property
    def field_domains(self):
        return self._field_domains

    @field_domains.setter
    def field_domains(self, domains):
        self._field_domains = domains

    @property
    def saved_fields(self):
        return self._saved_fields

    @property
    def source_power(self):
        return self._source_power

    @property
    def ndims(self):
        return self._ndims

    @abstractmethod
    def solve_forward(self): ...

    @abstractmethod
    def solve_adjoint(self): ...

    @abstractmethod
    def get_field(self, component, domain=None): ...

    @abstractmethod
    def get_field_interp(self, component, domain=None): ...

    @abstractmethod
    def get_adjoint_field(self, component, domain=None): ...

    @abstractmethod
    def build(self): ...

    @abstractmethod
    def update(self, bbox=None): ...

    @abstractmethod
    def set_sources(self, src, domain): ...

    @abstractmethod
    def set_adjoint_sources(self, src): ...

    @abstractmethod
    def update_saved_fields(self): ...

    @abstractmethod
    def get_source_power(self): ...

    @abstractmethod
    def get_A_diag(self): ...

    @abstractmethod
    def calc_ydAx(self, Adiag0): ...
