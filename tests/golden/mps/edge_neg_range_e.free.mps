NAME          EDGE_NEG_RANGE_E
ROWS
 N  COST
 E  C1
COLUMNS
    X1        COST           2.00000   C1           1.00000
RHS
    RHS1      C1            100.00000
RANGES
    RNG1      C1            -20.00000
BOUNDS
 UP BND       X1            200.00000
ENDATA
