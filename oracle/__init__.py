"""Oracle package: CPU restatement of the reference path.

TEST INFRASTRUCTURE. Importable only from tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as what results are compared against (and, in
that leg, as the timed CPU path). The product never imports it.
"""
