NAME          EDGE_FX_COST
ROWS
 N  COST    
 L  C1      
COLUMNS
    X1        COST           3.00000   C1             1.00000
    X2        COST           5.00000   C1             1.00000
RHS
    RHS1      C1            20.00000
BOUNDS
 FX BND       X1            10.00000
 UP BND       X2            50.00000
ENDATA
