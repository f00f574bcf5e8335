"""CPU oracle for SIGMo's hot path -- test infrastructure only (see oracle/oracle.cpp)."""
