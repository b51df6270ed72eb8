UCLA pl 1.0
c1 0 0 : N
c2 4 0 : N
c3 0 10 : N
m1 12 0 : N /FIXED
t1 8 0 : N /FIXED
p1 30 30 : N /FIXED_NI
