"""Model construction for the MI355X radiation-field library.

Restates the model-construction semantics of the reference Python package
(magritte/setup/: LAMDA line data, HEALPix rays, quadrature) and provides the
seeded synthetic models the library is measured on.
"""
