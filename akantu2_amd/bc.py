"""Boundary conditions with the reference's functors
(src/model/common/boundary_condition/boundary_condition_functor*.hh):

    model.applyBC(BC.Dirichlet.FixedValue(0.0, BC._x), "left")
    model.applyBC(BC.Neumann.FromTraction([0.0, 1e-3, 0.0]), "top")

Dirichlet functors act on the nodes of the group (blocked DOFs and the
displacement, boundary_condition_tmpl.hh:55-83).  FromTraction integrates a
constant traction over the group's boundary facets into the external force
on the device (akx_model_apply_traction, boundary_condition_tmpl.hh:100-160).
"""
from __future__ import annotations

import numpy as np

_x, _y, _z = 0, 1, 2


class Dirichlet:
    class FixedValue:
        """flags(axis) = true; u(axis) = value (boundary_condition_functor
        _inline_impl.hh:52-60)."""

        def __init__(self, value, axis):
            self.value, self.axis = float(value), int(axis)

        def apply(self, blocked, u, nodes):
            blocked[nodes, self.axis] = 1
            u[nodes, self.axis] = self.value

    class FlagOnly:
        """flags(axis) = true (boundary_condition_functor_inline_impl.hh:43-49)."""

        def __init__(self, axis):
            self.axis = int(axis)

        def apply(self, blocked, u, nodes):
            blocked[nodes, self.axis] = 1

    class IncrementValue:
        """flags(axis) = true; u(axis) += value (:62-70)."""

        def __init__(self, value, axis):
            self.value, self.axis = float(value), int(axis)

        def apply(self, blocked, u, nodes):
            blocked[nodes, self.axis] = 1
            u[nodes, self.axis] += self.value

    class Increment:
        """all flags true; u += value (:72-80)."""

        def __init__(self, value):
            self.value = np.asarray(value, dtype=np.float64)

        def apply(self, blocked, u, nodes):
            blocked[nodes, :] = 1
            u[nodes, :] += self.value


class Neumann:
    class FromTraction:
        """dual = traction on every facet quadrature point
        (boundary_condition_functor_inline_impl.hh:109-113)."""

        def __init__(self, traction):
            self.traction = np.asarray(traction, dtype=np.float64).reshape(3)
