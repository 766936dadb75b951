NAME	CRLF
* a comment line
ROWS
 N	obj

 L	r
COLUMNS
	x	obj	1	r	1
RHS
	rhs	r	2
ENDATA
