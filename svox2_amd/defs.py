"""Basis-type constants, with the values of the reference (svox2/defs.py:1-4,
mirroring svox2/csrc/include/data_spec.hpp:8-16).  Only BASIS_TYPE_SH is
rendered by this package; the other two are defined so that code written
against the reference imports unchanged, and are rejected at grid creation."""

BASIS_TYPE_SH = 1
BASIS_TYPE_3D_TEXTURE = 4
BASIS_TYPE_MLP = 255
