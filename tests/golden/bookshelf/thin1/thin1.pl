UCLA pl 1.0
# Created	:	Jan  6 2005
# User   	:	Gi-Joon Nam & Mehmet Yildiz at IBM Austin Research({gnam, mcan}@us.ibm.com)

o100851	0	0	: N
o100852	0	0	: N
o50015	0	0	: N
o211025	6240	22	: N /FIXED
