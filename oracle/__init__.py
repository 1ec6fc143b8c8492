"""ORACLE — test infrastructure only: the FP64 CPU restatement of the
reference's SIMPLE path (see oracle/oracle.hpp).  Not part of the product."""
