NAME          EDGE_DEGENERATE
ROWS
 N  COST
 L  C1
 L  C2
COLUMNS
    X1        COST           1.00000   C1           1.00000
    X2        COST           1.00000   C2           1.00000
RHS
    RHS1      C1             0.00000   C2             0.00000
ENDATA
