"""Backward-compatible module alias: ``import hmatrix_contact`` is ``aspher``."""
import sys

import aspher

sys.modules[__name__] = aspher
