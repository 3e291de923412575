"""Axisymmetric BEM geometries for the MI355X backend."""
