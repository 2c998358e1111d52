NAME          EDGE_MI_PL_BV
ROWS
 N  COST    
 L  C1      
COLUMNS
    MARK0000  'MARKER'                 'INTORG'
    X1        COST           1.00000   C1             1.00000
    X2        COST           2.00000   C1             1.00000
    MARK0001  'MARKER'                 'INTEND'
RHS
    RHS1      C1             5.00000
BOUNDS
 MI BND       X1      
 PL BND       X1      
 BV BND       X2      
ENDATA
