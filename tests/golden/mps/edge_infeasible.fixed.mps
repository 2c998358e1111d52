NAME          EDGE_INFEASIBLE
ROWS
 N  COST    
 L  C1      
 G  C2      
COLUMNS
    X1        COST           1.00000   C1             1.00000
    X1        C2             1.00000
RHS
    RHS1      C1             5.00000   C2            10.00000
ENDATA
