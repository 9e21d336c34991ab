NAME          TINY_SX
ROWS
 N  COST
 L  R1
 L  R2
COLUMNS
    X1        COST              -3.0
    X1        R1                 1.0
    X1        R2                 2.0
    X2        COST              -2.0
    X2        R1                 1.0
    X2        R2                 1.0
RHS
    RHS1      R1                 4.0
    RHS1      R2                 5.0
BOUNDS
 LO BND1      X1                 0.0
 LO BND1      X2                 0.0
ENDATA
