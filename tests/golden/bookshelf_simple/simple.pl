UCLA pl 1.0
# Created	:	Jan  6 2005
# User   	:	Gi-Joon Nam & Mehmet Yildiz at IBM Austin Research({gnam, mcan}@us.ibm.com)

o0 0 0 : N
o1 0 0 : N
o2 0 0 : N
o3 0 0 : N
o4 0 0 : N
o5 0 0 : N
o6 0 0 : N
o7 0 0 : N
o8 459 470 : N /Fixed
o9 500 555 : N /Fixed
